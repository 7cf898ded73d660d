// scan.hpp -- wave64 / block scans and reductions used by the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace thv {

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << lane_id()) - 1ull;
}

// The lanes where p holds (a 64-bit ballot of a bool).
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Set bits of m in the lanes below this one (two mbcnt instructions).
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Inclusive scan across the 64 lanes of a wave.
template <class T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const unsigned l = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T y = __shfl_up(x, d, 64);
    if (l >= (unsigned)d) x += y;
  }
  return x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// Block-wide exclusive scan of one value per thread; BS multiple of 64.
// `sh` must hold BS/64 + 1 elements.  Returns the exclusive prefix and sets
// *total.  Contains two barriers.
template <int BS, class T>
__device__ __forceinline__ T block_excl_scan(T x, T *sh, T *total) {
  constexpr int NW = BS / 64;
  const unsigned w = threadIdx.x / 64, l = lane_id();
  T inc = wave_incl_scan(x);
  if (l == 63) sh[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int i = 0; i < NW; i++) {
      T t = sh[i];
      sh[i] = run;
      run += t;
    }
    sh[NW] = run;
  }
  __syncthreads();
  T res = sh[w] + inc - x;
  *total = sh[NW];
  return res;
}

// Block-wide sum; `sh` holds BS/64 elements.  Result valid in all threads.
template <int BS, class T>
__device__ __forceinline__ T block_sum(T x, T *sh) {
  constexpr int NW = BS / 64;
  const unsigned w = threadIdx.x / 64, l = lane_id();
  x = wave_sum(x);
  __syncthreads();
  if (l == 0) sh[w] = x;
  __syncthreads();
  T r = 0;
#pragma unroll
  for (int i = 0; i < NW; i++) r += sh[i];
  return r;
}

}  // namespace thv
