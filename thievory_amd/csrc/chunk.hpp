// chunk.hpp -- the chunk machinery shared by every expansion kernel.
//
// A level's work space is the concatenation of the queued vertices' edge
// lists, one queue for host-resident and one for HBM-resident partitions.
// It is cut into chunks of C edges; a 256-thread block stages the queue
// entries of a chunk in LDS and every edge position then finds its entry by
// a popcount over an LDS bit mask of entry starts (see expand.hip).
#pragma once
#include "internal.hpp"
#include "scan.hpp"

namespace thv {

struct ChunkSrc {
  unsigned long long W[2], nq[2];
  uint32_t nch[2];
  uint32_t total;
};

__device__ __forceinline__ ChunkSrc chunk_src(const LevelCtr *ctr, uint32_t C) {
  ChunkSrc s;
  for (int q = 0; q < 2; q++) {
    s.W[q] = ctr->qw[q];
    s.nq[q] = ctr->qn[q];
    s.nch[q] = (uint32_t)((s.W[q] + C - 1) / C);
  }
  s.total = s.nch[Q_HOST] + s.nch[Q_DEV];
  return s;
}

// Combined chunk index -> (queue, chunk).  HBM chunks are spread evenly
// through the host chunks (Bresenham), so every block alternates between
// latency-bound zero-copy reads and HBM work.
__device__ __forceinline__ void chunk_pick(const ChunkSrc &s, uint32_t c, int &q, uint32_t &cc) {
  const unsigned long long D = s.nch[Q_DEV], T = s.total;
  const uint32_t d0 = (uint32_t)((unsigned long long)c * D / T);
  const uint32_t d1 = (uint32_t)((unsigned long long)(c + 1) * D / T);
  if (d1 > d0) {
    q = Q_DEV;
    cc = d0;
  } else {
    q = Q_HOST;
    cc = c - d0;
  }
}

struct ChunkInfo {
  unsigned long long e0;
  uint32_t len, i0, nent;
};

// Entries [i0, i0 + nent) intersect the chunk; entry i0 contains its first
// edge.  nent <= C + 1 because every queued vertex has at least one edge.
template <uint32_t C>
__device__ __forceinline__ ChunkInfo chunk_info(const ChunkSrc &s, const uint32_t *const *qc,
                                               int q, uint32_t cc) {
  ChunkInfo ci;
  ci.e0 = (unsigned long long)cc * C;
  ci.len = (uint32_t)min((unsigned long long)C, s.W[q] - ci.e0);
  ci.i0 = qc[q][cc];
  const uint32_t i1 = cc + 1 < s.nch[q] ? qc[q][cc + 1] : (uint32_t)(s.nq[q] - 1);
  ci.nent = i1 - ci.i0 + 1;
  return ci;
}

// Chunk-relative start of the entry whose global edge prefix is b (the
// first entry may start before the chunk); >= C means "starts in the next
// chunk" and the entry is not staged.
__device__ __forceinline__ unsigned long long chunk_rel(unsigned long long b,
                                                        unsigned long long e0) {
  return b > e0 ? b - e0 : 0;
}

// After every staged entry has set its start bit in mask[]: per-word
// exclusive popcount prefix, computed by the first wave.
template <uint32_t NW>
__device__ __forceinline__ void chunk_prefix(const uint32_t *mask, uint32_t *pre) {
  static_assert(NW <= 64, "one wave computes the prefix");
  if (threadIdx.x < 64) {
    const uint32_t t = threadIdx.x;
    const uint32_t pc = t < NW ? __popc(mask[t]) : 0u;
    const uint32_t inc = wave_incl_scan(pc);
    if (t < NW) pre[t] = inc - pc;
  }
}

// Entry of chunk position l (< len).
__device__ __forceinline__ uint32_t chunk_rank(const uint32_t *mask, const uint32_t *pre,
                                               uint32_t l) {
  const uint32_t wd = l >> 5;
  return pre[wd] + __popc(mask[wd] & (0xFFFFFFFFu >> (31 - (l & 31)))) - 1;
}

}  // namespace thv
