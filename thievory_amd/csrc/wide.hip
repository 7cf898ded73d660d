// wide.hip -- graphs with 64-bit vertex ids (2^32 - 1 vertices or more).
//
// The reference declares BFS64 / CC64 / SSSP64 for CSR<uint64> but leaves
// them empty (src/bfs/bfs.cu:544, src/cc/cc.cu:616, src/sssp/sssp.cu:572),
// and its file format stores 4-byte destinations (graph.cuh:754), so a
// graph past 2^32 vertices enters only through this API.  Such graphs are
// usually sparse in their id space, so the edge lists are kept by listed
// vertex: the sorted ids of the k vertices with edges, their list offsets
// (k + 1) and u64 destinations (with u32 weights for SSSP).  Values are
// dense: n x u64, with the reference's conventions (BFS 1 + hops, CC the
// least id reaching the vertex, SSSP min(n + 1, 1 + distance); unreached
// n + 1).
//
// Store: the listed vertices are cut into greedy runs of at most
// partition_edges edges (one oversize list alone, graph.cuh:270-304); the
// densest runs go to an HBM slab while the budget allows, the rest stay in
// mapped pinned host memory read zero-copy; eptr[j] is the address of
// listed vertex j's first record either way.
//
// Traversal: top-down, one wave per queued vertex (a 64-ary search of the
// sorted ids finds its list; lanes stride its records).  BFS claims a
// vertex with atomicCAS on its value and the winner appends it to the next
// queue; CC and SSSP lower values with atomicMin and append a vertex the
// first time it improves in a level (a bitmap marks queued vertices).
// Queues hold at most min(n, m + 1) ids.  Appends are wave-aggregated.
#include <algorithm>

#include "internal.hpp"
#include "scan.hpp"
#include "wide.hpp"

namespace thv {

namespace {

constexpr int WBS = 256;

// Index of v among the sorted listed ids (k of them), or k when v has no
// list: a 64-ary search, every lane loads one pivot per step.
__device__ __forceinline__ uint64_t find_listed(const uint64_t *ids, uint64_t k, uint64_t v) {
  const uint32_t lane = threadIdx.x & 63;
  uint64_t lo = 0, hi = k;  // answer in [lo, hi): the last index with ids[i] <= v
  while (hi - lo > 1) {
    const uint64_t step = (hi - lo + 63) / 64;
    const uint64_t p = lo + (uint64_t)lane * step;
    const bool le = p < hi && ids[p] <= v;
    const uint32_t c = (uint32_t)__popcll(__ballot(le));
    if (c == 0) return k;  // ids[lo] > v
    lo += (uint64_t)(c - 1) * step;
    hi = lo + step < hi ? lo + step : hi;
  }
  return (k > 0 && lo < k && ids[lo] == v) ? lo : k;
}

struct WideArgs {
  const uint64_t *ids, *loff, *eptr;
  uint64_t k, n;
  int rec;                      // 8: u64 dst; 16: {u64 dst, u32 w, pad}
  const uint64_t *q;            // this level's queue
  const unsigned long long *qn;
  uint64_t *nq;                 // next queue and its count
  unsigned long long *nqn;
  unsigned long long *edges;    // edges scanned (stats)
  uint64_t *val;
  uint32_t *inq;                // CC/SSSP: vertices in the next queue
  uint64_t level_value;         // BFS: value of the vertices this level finds
  int algo;
};

// Wave-aggregated append of the lanes with `take` set.
__device__ __forceinline__ void append(const WideArgs &a, bool take, uint64_t w) {
  const unsigned long long b = __ballot(take);
  if (!b) return;
  unsigned long long base = 0;
  if ((threadIdx.x & 63) == 0) base = atomicAdd(a.nqn, (unsigned long long)__popcll(b));
  base = __shfl(base, 0, 64);
  if (take) a.nq[base + __popcll(b & lanemask_lt())] = w;
}

__device__ __forceinline__ bool mark_queued(const WideArgs &a, uint64_t w) {
  const uint32_t bit = 1u << (w & 31);
  return !(atomicOr(&a.inq[w >> 5], bit) & bit);
}

// One wave per queued vertex (all k listed vertices when q is null: CC's
// first round).
__global__ void __launch_bounds__(WBS) k_wide_expand(WideArgs a) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nwaves = (uint64_t)gridDim.x * (WBS / 64);
  const uint64_t total = a.q ? *a.qn : a.k;
  unsigned long long scanned = 0;
  for (uint64_t qi = (uint64_t)blockIdx.x * (WBS / 64) + (threadIdx.x >> 6); qi < total;
       qi += nwaves) {
    uint64_t v, j;
    if (a.q) {
      v = a.q[qi];
      j = find_listed(a.ids, a.k, v);
      if (j == a.k) continue;
    } else {
      j = qi;
      v = a.ids[j];
    }
    const uint64_t d = a.loff[j + 1] - a.loff[j];
    const unsigned long long base = a.eptr[j];
    const uint64_t vv = a.val[v];
    scanned += lane == 0 ? d : 0;
    for (uint64_t e0 = 0; e0 < d; e0 += 64) {
      const uint64_t e = e0 + lane;
      bool take = false;
      uint64_t w = 0;
      if (e < d) {
        const unsigned char *r = reinterpret_cast<const unsigned char *>(base + e * a.rec);
        w = *reinterpret_cast<const uint64_t *>(r);
        if (a.algo == THV_ALGO_BFS) {
          const uint64_t none = a.n + 1;
          if (a.val[w] == none)
            take = atomicCAS((unsigned long long *)&a.val[w], none, a.level_value) == none;
        } else {
          const uint64_t nv =
              a.algo == THV_ALGO_SSSP ? vv + *reinterpret_cast<const uint32_t *>(r + 8) : vv;
          if (nv < a.val[w]) {
            const uint64_t old = atomicMin((unsigned long long *)&a.val[w], nv);
            take = nv < old && mark_queued(a, w);
          }
        }
      }
      append(a, take, w);
    }
  }
  scanned = wave_sum(scanned);
  if (lane == 0 && scanned) atomicAdd(a.edges, scanned);
}

// Clears the queued marks of this level's queue (they may re-enter).
__global__ void k_wide_unmark(const uint64_t *q, const unsigned long long *qn, uint32_t *inq) {
  const uint64_t n = *qn;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t w = q[i];
    atomicAnd(&inq[w >> 5], ~(1u << (w & 31)));
  }
}

__global__ void k_wide_fill(uint64_t *p, uint64_t v, uint64_t n, int iota) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = iota ? i : v;
}

__global__ void k_wide_set(uint64_t *p, uint64_t i, uint64_t v) { p[i] = v; }

// Digest of the values: sum over v with val[v] != none of
// mix64(mix64(v) + val[v]) (mod 2^64), the count and the largest value.
__global__ void __launch_bounds__(WBS) k_wide_digest(const uint64_t *val, uint64_t n,
                                                    uint64_t none, unsigned long long *out) {
  __shared__ unsigned long long sh[WBS / 64];
  unsigned long long h = 0, c = 0, mx = 0;
  for (uint64_t v = blockIdx.x * (uint64_t)WBS + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * WBS) {
    const uint64_t x = val[v];
    if (x != none) {
      h += mix64(mix64(v) + x);
      c++;
      mx = x > mx ? x : mx;
    }
  }
  const unsigned long long hs = block_sum<WBS>(h, sh), cs = block_sum<WBS>(c, sh);
  unsigned long long m = mx;
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long o = __shfl_xor(m, d, 64);
    m = o > m ? o : m;
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], hs);
    atomicAdd(&out[1], cs);
  }
  if ((threadIdx.x & 63) == 0) atomicMax(&out[2], m);
}

__global__ void k_wide_gather(const uint64_t *val, const uint64_t *at, uint64_t k,
                              uint64_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < k;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = val[at[i]];
}

int grid_of(uint64_t n, int cap) {
  return (int)std::max<uint64_t>(1, std::min<uint64_t>((n + WBS - 1) / WBS, (uint64_t)cap));
}

}  // namespace

Graph64::~Graph64() {
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
}

Graph64 *graph64_from_lists(uint64_t n, uint64_t k, const uint64_t *ids, const uint64_t *loff,
                            const uint64_t *dst, const uint32_t *w) {
  if (n == 0 || n == UINT64_MAX) fail(THV_EINVAL, "vertex count out of range");
  if (k > n) fail(THV_EINVAL, "more listed vertices than vertices");
  if (k && (!ids || !loff)) fail(THV_EINVAL, "null list arrays");
  if (k == 0 && loff && loff[0] != 0) fail(THV_EINVAL, "offsets must start at 0");
  const uint64_t m = k ? loff[k] : 0;
  if (m && !dst) fail(THV_EINVAL, "null destinations");
  if (k && loff[0] != 0) fail(THV_EINVAL, "offsets must start at 0");
  for (uint64_t j = 0; j < k; j++) {
    if (ids[j] >= n) fail(THV_EINVAL, "listed id out of range");
    if (j && ids[j] <= ids[j - 1]) fail(THV_EINVAL, "listed ids must be strictly ascending");
    if (loff[j + 1] <= loff[j]) fail(THV_EINVAL, "a listed vertex needs at least one edge");
  }
  for (uint64_t e = 0; e < m; e++)
    if (dst[e] >= n) fail(THV_EINVAL, "destination out of range");
  std::unique_ptr<Graph64> g(new Graph64());
  g->n = n;
  g->k = k;
  g->m = m;
  g->rec = w ? 16 : 8;
  g->h_ids.assign(ids, ids + k);
  g->h_loff.assign(loff, loff + k + 1);
  if (k == 0) g->h_loff.assign(1, 0);
  HIPC(hipGetDevice(&g->device));
  hipDeviceProp_t prop;
  HIPC(hipGetDeviceProperties(&prop, g->device));
  g->cus = prop.multiProcessorCount;
  HIPC(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  // records in mapped pinned host memory
  g->h_edges.alloc(std::max<uint64_t>(m, 1) * g->rec);
  unsigned char *r = static_cast<unsigned char *>(g->h_edges.p);
  for (uint64_t e = 0; e < m; e++) {
    memcpy(r + e * g->rec, &dst[e], 8);
    if (w) {
      memcpy(r + e * g->rec + 8, &w[e], 4);
      memset(r + e * g->rec + 12, 0, 4);
    }
  }
  g->d_ids.alloc(std::max<uint64_t>(k, 1) * 8);
  if (k) HIPC(hipMemcpy(g->d_ids.p, ids, k * 8, hipMemcpyHostToDevice));
  g->d_loff.alloc((k + 1) * 8);
  HIPC(hipMemcpy(g->d_loff.p, g->h_loff.data(), (k + 1) * 8, hipMemcpyHostToDevice));
  graph64_store(*g, UINT64_MAX, 1u << 16);
  return g.release();
}

void graph64_store(Graph64 &g, uint64_t budget, uint32_t part_edges) {
  if (part_edges == 0) fail(THV_EINVAL, "partition_edges must be > 0");
  HIPC(hipStreamSynchronize(g.stream));
  // greedy runs of listed vertices with at most part_edges edges
  std::vector<uint64_t> pj{0};
  uint64_t cur = 0;
  for (uint64_t j = 0; j < g.k; j++) {
    const uint64_t d = g.h_loff[j + 1] - g.h_loff[j];
    if (cur + d <= part_edges || j == pj.back()) {
      cur += d;
    } else {
      pj.push_back(j);
      cur = d;
    }
  }
  pj.push_back(g.k);
  const uint64_t P = pj.size() - 1;
  // densest runs first (edges per listed vertex), whole, while they fit
  std::vector<uint64_t> order(P);
  for (uint64_t p = 0; p < P; p++) order[p] = p;
  auto edges_of = [&](uint64_t p) { return g.h_loff[pj[p + 1]] - g.h_loff[pj[p]]; };
  auto dens = [&](uint64_t p) {
    return (double)edges_of(p) / (double)std::max<uint64_t>(1, pj[p + 1] - pj[p]);
  };
  std::stable_sort(order.begin(), order.end(),
                   [&](uint64_t a, uint64_t b) { return dens(a) > dens(b); });
  std::vector<uint8_t> res(P, 0);
  uint64_t used = 0;
  for (uint64_t p : order) {
    const uint64_t b = edges_of(p) * g.rec;
    if (budget != UINT64_MAX && used + b > budget) continue;
    res[p] = 1;
    used += b;
  }
  g.slab.alloc(std::max<uint64_t>(used, 8));
  std::vector<uint64_t> eptr(std::max<uint64_t>(g.k, 1));
  const unsigned char *host = static_cast<const unsigned char *>(g.h_edges.dev);
  uint64_t at = 0, dev_edges = 0;
  uint32_t pdev = 0;
  for (uint64_t p = 0; p < P; p++) {
    const uint64_t e0 = g.h_loff[pj[p]], ne = edges_of(p);
    if (res[p] && ne) {
      HIPC(hipMemcpy(static_cast<unsigned char *>(g.slab.p) + at,
                     static_cast<const unsigned char *>(g.h_edges.p) + e0 * g.rec, ne * g.rec,
                     hipMemcpyHostToDevice));
      for (uint64_t j = pj[p]; j < pj[p + 1]; j++)
        eptr[j] = (uint64_t)(static_cast<unsigned char *>(g.slab.p) + at +
                             (g.h_loff[j] - e0) * g.rec);
      at += ne * g.rec;
      dev_edges += ne;
      pdev++;
    } else {
      for (uint64_t j = pj[p]; j < pj[p + 1]; j++) eptr[j] = (uint64_t)(host + g.h_loff[j] * g.rec);
    }
  }
  g.d_eptr.alloc(eptr.size() * 8);
  HIPC(hipMemcpy(g.d_eptr.p, eptr.data(), eptr.size() * 8, hipMemcpyHostToDevice));
  g.P = (uint32_t)P;
  g.Pdev = pdev;
  g.dev_edges = dev_edges;
  g.budget = budget;
}

static void ensure_work(Graph64 &g, bool marks) {
  const uint64_t qcap = std::min<uint64_t>(g.n, g.m + 1);
  if (!g.val.p) g.val.alloc(g.n * 8);
  if (g.q[0].bytes < qcap * 8)
    for (auto &q : g.q) q.alloc(qcap * 8);
  if (!g.ctr.p) g.ctr.alloc(8 * 8);
  if (marks && !g.inq.p) {
    const uint64_t words = g.n / 32 + 2;
    g.inq.alloc(words * 4);
    HIPC(hipMemsetAsync(g.inq.p, 0, words * 4, g.stream));
  }
}

void run_wide(Graph64 &g, int algo, uint64_t src, uint64_t *out, thv_stats *st,
              uint64_t *digest3) {
  if (algo != THV_ALGO_CC && src >= g.n) fail(THV_EINVAL, "source vertex out of range");
  if (algo == THV_ALGO_SSSP && g.rec != 16)
    fail(THV_ENOTSUP, "SSSP needs weights (build the graph with them)");
  ensure_work(g, algo != THV_ALGO_BFS);
  hipStream_t s = g.stream;
  const uint64_t none = g.n + 1;
  unsigned long long *ctr = g.ctr.as<unsigned long long>();  // [0],[1] counts, [2] edges
  if (st) memset(st, 0, sizeof *st);
  HIPC(hipMemsetAsync(ctr, 0, 8 * 8, s));
  const int fg = grid_of(g.n, 16 * g.cus);
  if (algo == THV_ALGO_CC) {
    hipLaunchKernelGGL(k_wide_fill, dim3(fg), dim3(WBS), 0, s, g.val.as<uint64_t>(), 0ull, g.n, 1);
  } else {
    hipLaunchKernelGGL(k_wide_fill, dim3(fg), dim3(WBS), 0, s, g.val.as<uint64_t>(), none, g.n, 0);
    hipLaunchKernelGGL(k_wide_set, dim3(1), dim3(1), 0, s, g.val.as<uint64_t>(), src, 1ull);
    hipLaunchKernelGGL(k_wide_set, dim3(1), dim3(1), 0, s, g.q[0].as<uint64_t>(), 0ull, src);
    hipLaunchKernelGGL(k_wide_set, dim3(1), dim3(1), 0, s, (uint64_t *)ctr, 0ull, 1ull);
  }
  HIPC(hipStreamSynchronize(s));
  WideArgs a;
  memset(&a, 0, sizeof a);
  a.ids = g.d_ids.as<uint64_t>();
  a.loff = g.d_loff.as<uint64_t>();
  a.eptr = g.d_eptr.as<uint64_t>();
  a.k = g.k;
  a.n = g.n;
  a.rec = g.rec;
  a.val = g.val.as<uint64_t>();
  a.inq = algo == THV_ALGO_BFS ? nullptr : g.inq.as<uint32_t>();
  a.edges = ctr + 2;
  a.algo = algo;
  const int eg = 8 * g.cus;
  const uint64_t t0 = now_ns();
  int cur = 0;
  uint64_t levels = 0;
  for (uint64_t L = 0;; L++) {
    // CC's first round runs over every listed vertex; later rounds, and
    // every BFS / SSSP round, over the queue
    const bool all = algo == THV_ALGO_CC && L == 0;
    a.q = all ? nullptr : g.q[cur].as<uint64_t>();
    a.qn = ctr + cur;
    a.nq = g.q[cur ^ 1].as<uint64_t>();
    a.nqn = ctr + (cur ^ 1);
    a.level_value = L + 2;
    HIPC(hipMemsetAsync(a.nqn, 0, 8, s));
    if (a.inq && !all)
      hipLaunchKernelGGL(k_wide_unmark, dim3(fg), dim3(WBS), 0, s, a.q, a.qn, a.inq);
    hipLaunchKernelGGL(k_wide_expand, dim3(eg), dim3(WBS), 0, s, a);
    unsigned long long next = 0;
    HIPC(hipMemcpyAsync(&next, a.nqn, 8, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    levels++;
    if (next == 0) break;
    cur ^= 1;
  }
  if (a.inq) {  // leave the marks clear for the next call
    a.q = g.q[cur ^ 1].as<uint64_t>();
    a.qn = ctr + (cur ^ 1);
    hipLaunchKernelGGL(k_wide_unmark, dim3(fg), dim3(WBS), 0, s, a.q, a.qn, a.inq);
  }
  HIPC(hipStreamSynchronize(s));
  const double ms = (now_ns() - t0) * 1e-6;
  unsigned long long edges = 0;
  HIPC(hipMemcpy(&edges, ctr + 2, 8, hipMemcpyDeviceToHost));
  uint64_t dg[3] = {0, 0, 0};
  if (digest3 || st) {
    HIPC(hipMemsetAsync(ctr + 4, 0, 3 * 8, s));
    hipLaunchKernelGGL(k_wide_digest, dim3(fg), dim3(WBS), 0, s, (const uint64_t *)g.val.p, g.n,
                       algo == THV_ALGO_CC ? UINT64_MAX : none, ctr + 4);
    HIPC(hipMemcpyAsync(dg, ctr + 4, 3 * 8, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
  }
  if (digest3) memcpy(digest3, dg, sizeof dg);
  if (out) HIPC(hipMemcpy(out, g.val.p, g.n * 8, hipMemcpyDeviceToHost));
  if (st) {
    st->ms = ms;
    st->iterations = (uint32_t)levels;
    st->edges_traversed = edges;
    st->edges_examined = edges;
    st->reached = dg[1];
    st->expand_launches = (uint32_t)levels;
  }
}

void wide_gather(Graph64 &g, const uint64_t *at, uint64_t k, uint64_t *out) {
  if (!g.val.p) fail(THV_ESTATE, "no values yet: run an algorithm first");
  for (uint64_t i = 0; i < k; i++)
    if (at[i] >= g.n) fail(THV_EINVAL, "vertex out of range");
  if (k == 0) return;
  DevMem d_at, d_out;
  d_at.alloc(k * 8);
  d_out.alloc(k * 8);
  HIPC(hipMemcpy(d_at.p, at, k * 8, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_wide_gather, dim3(grid_of(k, 1024)), dim3(WBS), 0, g.stream,
                     (const uint64_t *)g.val.p, (const uint64_t *)d_at.p, k, (uint64_t *)d_out.p);
  HIPC(hipStreamSynchronize(g.stream));
  HIPC(hipMemcpy(out, d_out.p, k * 8, hipMemcpyDeviceToHost));
}

}  // namespace thv
