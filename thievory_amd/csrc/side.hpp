// side.hpp -- the device halves of LevelSide: a grid's share of the words a
// launch clears for the next level, and the copy of a level's finished
// counters into the host-mapped readback.
#pragma once
#include "internal.hpp"

namespace thv {

static_assert(sizeof(LevelCtr) % 8 == 0, "LevelCtr is copied as u64 words");

// Clears s.zero[0, s.nzero) across the whole grid: 16-byte stores over the
// (256-byte aligned) array, then the tail words.
__device__ __forceinline__ void side_clear(const LevelSide &s) {
  if (!s.zero) return;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
  uint4 *z4 = reinterpret_cast<uint4 *>(s.zero);
  const uint64_t n4 = s.nzero / 4;
  for (uint64_t i = t; i < n4; i += T) z4[i] = make_uint4(0u, 0u, 0u, 0u);
  for (uint64_t i = n4 * 4 + t; i < s.nzero; i += T) s.zero[i] = 0u;
}

// Run by the 64 lanes of one wave once every counter it copies is final:
// written by an earlier launch, or (own_at) the launch's own LevelCtr, whose
// value `own` the caller passes (from registers or LDS, never re-read from
// memory this launch wrote).  Lane i copies u64 word i to the host-mapped
// readback, the extra counters likewise; then zctr is reset, each of its
// words by the lane that copied it (so after that lane's own load).
__device__ __forceinline__ void side_publish(const LevelSide &s, uint32_t lane,
                                             const LevelCtr *own_at = nullptr,
                                             const LevelCtr *own = nullptr) {
  constexpr uint32_t W = (uint32_t)(sizeof(LevelCtr) / 8);
  const bool zin = s.pub && s.zctr >= s.pub_src && s.zctr < s.pub_src + s.npub;
  if (s.pub && lane < s.npub * W) {
    const uint32_t slot = lane / W, k = lane % W;
    const LevelCtr *c = s.pub_src + slot;
    unsigned long long *src =
        const_cast<unsigned long long *>(reinterpret_cast<const unsigned long long *>(c)) + k;
    const unsigned long long v =
        own && c == own_at ? reinterpret_cast<const unsigned long long *>(own)[k] : *src;
    reinterpret_cast<unsigned long long *>(s.pub + slot)[k] = v;
    if (zin && c == s.zctr) *src = 0ull;
  }
  if (s.xpub && lane < s.nx) s.xpub[lane] = s.xsrc[lane];
  if (s.zctr && !zin && lane < W) reinterpret_cast<unsigned long long *>(s.zctr)[lane] = 0ull;
  __threadfence_system();
}

}  // namespace thv
