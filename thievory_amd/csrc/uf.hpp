// uf.hpp -- union-find on the GPU for connected components of a symmetric
// graph.  comp[x] <= x always holds and a link hooks the larger root under
// the smaller one, so every root is the smallest vertex id of its tree and
// the final labels are the component minima -- the same labels as min-label
// propagation (the reference's CC, cc_kernels.cu:50-66) on a symmetric
// graph.
//
// Loads of comp[] are device-coherent atomics, so a retry always sees the
// parents other waves wrote.  Each retry lowers the larger of the two roots
// it works on, so a link ends after at most that many steps; the bound
// max_steps (THV_OPT_CC_UF_STEPS, default 2^20) is a guard that
// raises *err instead of looping on, and the host then falls back to label
// propagation.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace thv {

__device__ __forceinline__ uint32_t uf_load(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void uf_link(uint32_t u, uint32_t v, uint32_t *comp, uint32_t *err,
                                        uint32_t max_steps) {
  uint32_t p1 = uf_load(&comp[u]), p2 = uf_load(&comp[v]);
  for (uint32_t step = 0; p1 != p2; step++) {
    if (step >= max_steps) {
      atomicOr(err, 1u);
      return;
    }
    const uint32_t high = p1 > p2 ? p1 : p2, low = p1 > p2 ? p2 : p1;
    const uint32_t ph = uf_load(&comp[high]);
    if (ph == low) return;
    if (ph == high && atomicCAS(&comp[high], high, low) == high) return;
    p1 = uf_load(&comp[uf_load(&comp[high])]);
    p2 = uf_load(&comp[low]);
  }
}

// Root of x with path halving left to the caller's compress pass.
__device__ __forceinline__ uint32_t uf_find(const uint32_t *comp, uint32_t x, uint32_t *err,
                                            uint32_t max_steps) {
  uint32_t c = uf_load(&comp[x]);
  for (uint32_t step = 0;; step++) {
    const uint32_t p = uf_load(&comp[c]);
    if (p == c) return c;
    if (step >= max_steps) {
      atomicOr(err, 1u);
      return c;
    }
    c = p;
  }
}

}  // namespace thv
