// cc.hip -- connected components of a symmetric graph by union-find with
// neighbour sampling (Afforest: Sutton, Ben-Nun, Bader, IPDPS 2018).
//
// The reference computes CC by min-label propagation along out-edges until
// no label changes (cc/cc.cu:71-614, cc_kernels.cu:50-66); on a symmetric
// graph its fixpoint labels every vertex with the smallest id of its
// component.  Union-find with "larger root under smaller" hooking (uf.hpp)
// ends with exactly those labels, and reads far fewer edges:
//   1. link every vertex to its first two neighbours, compress (a vertex
//      whose list is on the host links its parent hint instead, when the
//      store holds hints: one HBM word, no request over the link);
//   2. sample 1024 vertices: the most frequent root is the giant component;
//   3. link the remaining lists of the vertices outside it (its own
//      members need nothing more: an edge to a non-member is linked from
//      the other side, which the graph's symmetry guarantees; a hinted
//      host list counts as remaining whatever its length), compress.
// Directed graphs keep label propagation (the labels there are not
// components).
#include <algorithm>
#include <unordered_map>

#include "internal.hpp"
#include "scan.hpp"
#include "uf.hpp"

namespace thv {

namespace {

constexpr int UBS = 256;
constexpr uint32_t CC_SAMPLES = 1024;

// Step 1, one vertex per thread: its first two records, then the links;
// with hints (nullable), a host list's vertex links its hint instead.
template <int REC>
__global__ void __launch_bounds__(UBS)
k_cc_first(const uint64_t *eptr, const uint32_t *deg, const uint32_t *rbit, const uint32_t *hint,
           uint64_t n, uint32_t *comp, uint32_t *err, uint32_t steps, unsigned long long *cnt) {
  __shared__ unsigned long long sh[UBS / 64];
  unsigned long long rd[2] = {0, 0};  // records read (host, HBM); a hint counts as HBM
  for (uint64_t v = blockIdx.x * (uint64_t)UBS + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * UBS) {
    const uint32_t d = deg[v];
    if (!d) continue;
    if (hint && !((rbit[v >> 5] >> (v & 31)) & 1u)) {
      uf_link((uint32_t)v, hint[v], comp, err, steps);  // d > 0: a neighbour
      rd[1] += 1;
      continue;
    }
    const unsigned long long b = eptr[v];
    const uint32_t w0 = *reinterpret_cast<const uint32_t *>(b);
    const uint32_t w1 = d > 1 ? *reinterpret_cast<const uint32_t *>(b + REC) : w0;
    uf_link((uint32_t)v, w0, comp, err, steps);
    if (d > 1) uf_link((uint32_t)v, w1, comp, err, steps);
    rd[(rbit[v >> 5] >> (v & 31)) & 1u] += d > 1 ? 2 : 1;
  }
  const unsigned long long h = block_sum<UBS>(rd[0], sh);
  const unsigned long long m = block_sum<UBS>(rd[1], sh);
  if (threadIdx.x == 0) {
    if (h) atomicAdd(&cnt[0], h);
    if (m) atomicAdd(&cnt[1], m);
  }
}

__global__ void __launch_bounds__(UBS) k_cc_compress(uint32_t *comp, uint64_t n, uint32_t *err,
                                                     uint32_t steps) {
  for (uint64_t v = blockIdx.x * (uint64_t)UBS + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * UBS)
    comp[v] = uf_find(comp, (uint32_t)v, err, steps);
}

__global__ void k_cc_sample(const uint32_t *comp, uint64_t n, uint32_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < CC_SAMPLES) out[i] = comp[mix64(i) % n];
}

// Vertices whose remaining lists are linked in step 3: a root other than
// the giant's, and more than two records or (hinted) a host list.  One lane
// per vertex; lanes 0 and 32 write the wave's two bitmap words.
__global__ void __launch_bounds__(UBS)
k_cc_rest(const uint32_t *comp, const uint32_t *deg, const uint32_t *rbit, bool hinted,
          uint64_t n, uint32_t giant, uint32_t *bits) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t v = blockIdx.x * (uint64_t)UBS + threadIdx.x; v - lane < n;
       v += (uint64_t)gridDim.x * UBS) {
    bool sel = false;
    if (v < n) {
      const uint32_t d = deg[v];
      const bool rest = d > 2 || (hinted && d > 0 && !((rbit[v >> 5] >> (v & 31)) & 1u));
      sel = rest && comp[v] != giant;
    }
    const unsigned long long b = __ballot(sel);
    if (lane == 0 || lane == 32) {
      const uint64_t w = v >> 5;
      if (w < (n + 31) / 32) bits[w] = (uint32_t)(b >> lane);
    }
  }
}

int grid_for(uint64_t n, int cap) {
  return (int)std::max<uint64_t>(1, std::min<uint64_t>((n + UBS - 1) / UBS, (uint64_t)cap));
}

}  // namespace

bool cc_union_find(Graph &g, thv_stats *st) {
  Workspace &w = g.ws;
  hipStream_t s = g.stream;
  const uint64_t n = g.n, words = bitmap_words(n);
  uint32_t *comp = w.val.as<uint32_t>();
  unsigned long long *aux = w.aux.as<unsigned long long>();
  uint32_t *err = reinterpret_cast<uint32_t *>(aux + 2);
  HIPC(hipMemsetAsync(aux, 0, 3 * 8, s));  // [0], [1] records read, [2] error flag
  const uint64_t t0 = now_ns();
  const int grid = grid_for(n, 16 * g.cus);
  const uint32_t steps = g.uf_max_steps;
  const uint64_t *eptr = g.store.d_eptr.as<uint64_t>();
  const uint32_t *deg = g.d_deg.as<uint32_t>(), *rbit = g.store.d_rbit.as<uint32_t>();
  const uint32_t *hint = g.store.d_hint.as<uint32_t>();  // null without hints
  if (g.rec == 8)
    hipLaunchKernelGGL(k_cc_first<8>, dim3(grid), dim3(UBS), 0, s, eptr, deg, rbit, hint, n, comp,
                       err, steps, aux);
  else
    hipLaunchKernelGGL(k_cc_first<4>, dim3(grid), dim3(UBS), 0, s, eptr, deg, rbit, hint, n, comp,
                       err, steps, aux);
  hipLaunchKernelGGL(k_cc_compress, dim3(grid), dim3(UBS), 0, s, comp, n, err, steps);
  // the sample lands in the probe's short-list buffer (unused here; at
  // least one tile = 16 KB, CC_SAMPLES x 4 = 4 KB), then on the host
  static_assert(CC_SAMPLES * 4 <= TILE * 4, "sample buffer");
  if (w.qs.slist.bytes < CC_SAMPLES * 4) fail(THV_ESTATE, "CC sample buffer too small");
  uint32_t *samp = w.qs.slist.as<uint32_t>();
  hipLaunchKernelGGL(k_cc_sample, dim3(CC_SAMPLES / UBS), dim3(UBS), 0, s, comp, n, samp);
  std::vector<uint32_t> hs(CC_SAMPLES);
  HIPC(hipMemcpyAsync(hs.data(), samp, CC_SAMPLES * 4, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  std::unordered_map<uint32_t, uint32_t> freq;
  uint32_t giant = hs[0], best = 0;
  for (uint32_t c : hs)
    if (++freq[c] > best) {
      best = freq[c];
      giant = c;
    }
  // step 3: the rest of the lists outside the giant component
  uint32_t *bits = w.front[0].as<uint32_t>();
  HIPC(hipMemsetAsync(bits, 0, words * 4, s));
  hipLaunchKernelGGL(k_cc_rest, dim3(grid), dim3(UBS), 0, s, comp, deg, rbit, hint != nullptr, n,
                     giant, bits);
  LevelCtr *c = w.qs.ctr.as<LevelCtr>();
  launch_compact(g, w.qs, bits, c, s);
  ExpandArgs a = expand_args(g, w.qs);
  a.ctr = c;
  a.val = comp;
  a.err = err;
  a.uf_steps = steps;
  HIPC(hipEventRecord(w.ev[0], s));
  launch_expand(g, A_CC_LINK, a, s);
  HIPC(hipEventRecord(w.ev[1], s));
  hipLaunchKernelGGL(k_cc_compress, dim3(grid), dim3(UBS), 0, s, comp, n, err, steps);
  HIPC(hipMemcpyAsync(w.h_aux, aux, 3 * 8, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(&w.h_ctr[0], c, sizeof(LevelCtr), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (w.h_aux[2]) return false;  // a link hit its step guard: not trusted
  if (st) {
    const LevelCtr &h = w.h_ctr[0];
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
    const uint64_t rec = (uint64_t)g.rec;
    const uint64_t first = w.h_aux[0] + w.h_aux[1];
    st->edges_traversed = first + h.qw[0] + h.qw[1];
    st->edges_examined = st->edges_traversed;
    st->host_bytes = (w.h_aux[0] + h.qw[0]) * rec;
    // records in HBM, 8 bytes of comp traffic per linked record, the degree
    // and first address of every vertex and comp twice per compress pass
    st->hbm_bytes = (w.h_aux[1] + h.qw[1]) * rec + 8 * st->edges_traversed + 12 * n + 16 * n;
    st->iterations = 2;
    st->expand_launches = 1;
    st->expand_ms = ms;
    st->ms = (now_ns() - t0) * 1e-6;
  }
  return true;
}

}  // namespace thv
