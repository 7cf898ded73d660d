// rmat.hip -- seeded R-MAT graph generation on the GPU.
//
// No reference counterpart (the reference only reads files); the synthetic
// inputs of every benchmark configuration are built here.  The tuple stream
// is the same counter-based one the CPU generator in the oracle uses
// (SplitMix64, integer quadrant thresholds), so both produce bit-identical
// CSR arrays: tuples -> self-loops dropped -> (optionally) both directions
// -> 64-bit (src << 32 | dst) keys radix-sorted on the GPU -> duplicates
// dropped -> offsets from a degree histogram -> symmetric weights in
// [1, 64] from a hash of the unordered pair.
#include <cmath>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "internal.hpp"

namespace thv {

struct RmatParams {
  int scale;
  uint64_t seed;
  uint64_t th0, th1, th2;
};

__device__ __forceinline__ void rmat_tuple(const RmatParams &p, uint64_t e, uint32_t &u,
                                           uint32_t &v) {
  const uint64_t key = mix64(p.seed * 0x9E3779B97F4A7C15ull + e);
  uint32_t uu = 0, vv = 0;
  for (int l = 0; l < p.scale; l++) {
    const uint64_t r = mix64(key + (uint64_t)l) >> 32;
    uint32_t bu, bv;
    if (r < p.th0) { bu = 0; bv = 0; }
    else if (r < p.th1) { bu = 0; bv = 1; }
    else if (r < p.th2) { bu = 1; bv = 0; }
    else { bu = 1; bv = 1; }
    uu = (uu << 1) | bu;
    vv = (vv << 1) | bv;
  }
  u = uu;
  v = vv;
}

// keys[] receive (row << 32 | col) with row = src (CSR) or dst (transpose);
// self-loops become `invalid`, which sorts after every real key.
__global__ void k_tuples(RmatParams p, uint64_t ntup, bool symmetric, bool transpose,
                         uint64_t invalid, uint64_t *keys) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < ntup;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u, v;
    rmat_tuple(p, e, u, v);
    const bool ok = u != v;
    if (symmetric) {
      keys[2 * e] = ok ? ((uint64_t)u << 32 | v) : invalid;
      keys[2 * e + 1] = ok ? ((uint64_t)v << 32 | u) : invalid;
    } else {
      const uint64_t k = transpose ? ((uint64_t)v << 32 | u) : ((uint64_t)u << 32 | v);
      keys[e] = ok ? k : invalid;
    }
  }
}

__global__ void k_flags(const uint64_t *keys, uint64_t cnt, uint64_t invalid, uint32_t *flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    flag[i] = (k != invalid && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
  }
}

__host__ __device__ __forceinline__ uint32_t edge_weight(uint64_t seed, uint32_t u, uint32_t v) {
  const uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
  return 1u + (uint32_t)(mix64(seed * 0xD1B54A32D192ED03ull + ((lo << 32) | hi)) & 63u);
}

// Scatter unique keys to their compacted positions; count row lengths and,
// for a transpose, the out-degree of every column vertex.
__global__ void k_scatter(const uint64_t *keys, const uint32_t *flag, const uint32_t *pos,
                          uint64_t cnt, int rec, bool weights, uint64_t wseed, void *records,
                          unsigned int *rowdeg, unsigned int *coldeg) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (!flag[i]) continue;
    const uint64_t k = keys[i];
    const uint32_t r = (uint32_t)(k >> 32), c = (uint32_t)k;
    const uint64_t p = pos[i];
    if (rec == 8) {
      uint2 x;
      x.x = c;
      x.y = weights ? edge_weight(wseed, r, c) : 20u;
      reinterpret_cast<uint2 *>(records)[p] = x;
    } else {
      reinterpret_cast<uint32_t *>(records)[p] = c;
    }
    atomicAdd(&rowdeg[r], 1u);
    if (coldeg) atomicAdd(&coldeg[c], 1u);
  }
}

static int grid_for(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(b, 16384));
}

Graph *rmat_generate(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                     bool symmetric, bool weights, bool transpose) {
  if (scale < 1 || scale > 31) fail(THV_EINVAL, "scale must be in [1, 31]");
  if (edge_factor < 1) fail(THV_EINVAL, "edge_factor must be >= 1");
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) fail(THV_EINVAL, "bad R-MAT probabilities");
  if (symmetric && transpose) fail(THV_EINVAL, "a symmetric graph is its own transpose");
  const uint64_t n = 1ull << scale;
  const uint64_t ntup = symmetric ? (uint64_t)edge_factor * n / 2 : (uint64_t)edge_factor * n;
  const uint64_t cnt = symmetric ? 2 * ntup : ntup;
  if (cnt >= 0xFFFFFFFFull) fail(THV_ENOTSUP, "more than 2^32-1 directed edges");

  std::unique_ptr<Graph> g(new Graph());
  graph_init_device(*g);
  hipStream_t s = g->stream;
  g->n = n;
  g->rec = weights ? 8 : 4;
  g->csc = transpose;
  g->sym = symmetric;

  RmatParams p;
  p.scale = scale;
  p.seed = seed;
  p.th0 = (uint64_t)floor(a * 4294967296.0);
  p.th1 = (uint64_t)floor((a + b) * 4294967296.0);
  p.th2 = (uint64_t)floor((a + b + c) * 4294967296.0);
  const uint64_t invalid = ((n - 1) << 32) | 0xFFFFFFFFull;

  DevMem kA, kB;
  kA.alloc(cnt * 8);
  kB.alloc(cnt * 8);
  hipLaunchKernelGGL(k_tuples, dim3(grid_for(ntup)), dim3(256), 0, s, p, ntup, symmetric,
                     transpose, invalid, kA.as<uint64_t>());
  HIPC(hipGetLastError());

  rocprim::double_buffer<uint64_t> db(kA.as<uint64_t>(), kB.as<uint64_t>());
  size_t tmp_bytes = 0;
  HIPC(rocprim::radix_sort_keys(nullptr, tmp_bytes, db, (size_t)cnt, 0u,
                                (unsigned)(32 + scale), s));
  DevMem tmp;
  tmp.alloc(tmp_bytes);
  HIPC(rocprim::radix_sort_keys(tmp.p, tmp_bytes, db, (size_t)cnt, 0u, (unsigned)(32 + scale),
                                s));
  uint64_t *keys = db.current();
  // the other buffer holds flags (first half) and positions (second half)
  uint32_t *flag = reinterpret_cast<uint32_t *>(db.alternate());
  uint32_t *pos = flag + cnt;
  hipLaunchKernelGGL(k_flags, dim3(grid_for(cnt)), dim3(256), 0, s, keys, cnt, invalid, flag);
  size_t sbytes = 0;
  HIPC(rocprim::exclusive_scan(nullptr, sbytes, flag, pos, 0u, (size_t)cnt,
                               rocprim::plus<uint32_t>(), s));
  if (sbytes > tmp.bytes) tmp.alloc(sbytes);
  HIPC(rocprim::exclusive_scan(tmp.p, sbytes, flag, pos, 0u, (size_t)cnt,
                               rocprim::plus<uint32_t>(), s));
  uint32_t last_pos = 0, last_flag = 0;
  HIPC(hipMemcpyAsync(&last_pos, pos + cnt - 1, 4, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(&last_flag, flag + cnt - 1, 4, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  const uint64_t m = (uint64_t)last_pos + last_flag;
  g->m = m;

  DevMem recs, rowdeg, coldeg, off;
  recs.alloc(m * g->rec);
  rowdeg.alloc(n * 4);
  HIPC(hipMemsetAsync(rowdeg.p, 0, n * 4, s));
  if (transpose) {
    coldeg.alloc(n * 4);
    HIPC(hipMemsetAsync(coldeg.p, 0, n * 4, s));
  }
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(cnt)), dim3(256), 0, s, keys, flag, pos, cnt,
                     g->rec, weights, seed, recs.p, rowdeg.as<unsigned int>(),
                     transpose ? coldeg.as<unsigned int>() : (unsigned int *)nullptr);
  HIPC(hipGetLastError());
  tmp.release();
  kB.release();
  kA.release();

  // offsets = exclusive scan of row degrees (n + 1 entries)
  off.alloc((n + 1) * 8);
  size_t obytes = 0;
  auto degs = rocprim::make_transform_iterator(rowdeg.as<unsigned int>(),
                                               [] __device__(unsigned int d) {
                                                 return (unsigned long long)d;
                                               });
  HIPC(rocprim::exclusive_scan(nullptr, obytes, degs, off.as<unsigned long long>(), 0ull,
                               (size_t)n, rocprim::plus<unsigned long long>(), s));
  tmp.alloc(obytes);
  HIPC(rocprim::exclusive_scan(tmp.p, obytes, degs, off.as<unsigned long long>(), 0ull,
                               (size_t)n, rocprim::plus<unsigned long long>(), s));
  g->h_off.resize(n + 1);
  HIPC(hipMemcpyAsync(g->h_off.data(), off.p, n * 8, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  g->h_off[n] = m;

  g->h_edges.alloc(m * g->rec);
  HIPC(hipMemcpy(g->h_edges.p, recs.p, m * g->rec, hipMemcpyDeviceToHost));
  if (transpose) {
    g->d_outdeg.alloc(n * 4);
    HIPC(hipMemcpy(g->d_outdeg.p, coldeg.p, n * 4, hipMemcpyDeviceToDevice));
  }
  recs.release();
  graph_build_tables(*g);
  return g.release();
}

}  // namespace thv
