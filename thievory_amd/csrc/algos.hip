// algos.hip -- host iteration loops of BFS, CC, SSSP and PageRank.
//
// Reference drivers: BFS32 (src/bfs/bfs.cu:8-542), CC32 (src/cc/cc.cu:
// 71-614), SSSP32 (src/sssp/sssp.cu:10-570), PR32 pull (src/pr/pr.cu:
// 11-519) and PR32_PUSH (src/pr/pr.cu:1030-1549).  Each reference level does
// about 8 O(n) passes and 6-10 blocking host syncs; here a level is
// compaction (3 kernels) + expansion (1 kernel) + one 48-byte readback.
// Initial values follow CSR::InitFrontierNValues (src/graph.cuh:229-268).
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "internal.hpp"
#include "scan.hpp"

namespace thv {

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_iota_u32(uint32_t *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

// Bitmap with bits [0, n) set and every later bit of the `words` words clear.
__global__ void k_bits_prefix(uint32_t *bm, uint64_t n, uint64_t words) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = i * 32;
    uint32_t x = 0;
    if (lo + 32 <= n) x = 0xFFFFFFFFu;
    else if (lo < n) x = (1u << (n - lo)) - 1u;
    bm[i] = x;
  }
}

__global__ void k_set_bit(uint32_t *bm, uint32_t v) { bm[v >> 5] |= 1u << (v & 31); }

// The start of a BFS from src in one launch: every value n + 1 but src's 1
// (16-byte stores over the 256-byte aligned array); visited = the degree-0
// vertices (iso, may be null) and src; frontier f0 = {src}; f1 and f2 (may
// be null) clear; the u64 counters z0[0, nz0) and z1[0, nz1) zero.
__global__ void __launch_bounds__(256)
k_bfs_init(uint32_t *val, uint64_t n, uint32_t src, uint32_t *vis, const uint32_t *iso,
           uint32_t *f0, uint32_t *f1, uint32_t *f2, uint64_t words, unsigned long long *z0,
           uint32_t nz0, unsigned long long *z1, uint32_t nz1) {
  const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
  const uint32_t far = (uint32_t)(n + 1);
  uint4 *v4 = reinterpret_cast<uint4 *>(val);
  for (uint64_t i = t; i < n / 4; i += T) {
    uint4 x = make_uint4(far, far, far, far);
    if (i == src / 4) {
      const uint32_t k = src & 3u;
      if (k == 0) x.x = 1u;
      else if (k == 1) x.y = 1u;
      else if (k == 2) x.z = 1u;
      else x.w = 1u;
    }
    // a streaming store: no level reads these lines soon (measured at C2:
    // the start 2% faster per BFS than with ordinary stores)
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 y = {x.x, x.y, x.z, x.w};
    __builtin_nontemporal_store(y, reinterpret_cast<u32x4 *>(&v4[i]));
  }
  for (uint64_t i = n / 4 * 4 + t; i < n; i += T) val[i] = i == src ? 1u : far;
  const uint64_t sw = src >> 5;
  const uint32_t sb = 1u << (src & 31);
  for (uint64_t i = t; i < words; i += T) {
    const uint32_t b = i == sw ? sb : 0u;
    vis[i] = (iso ? iso[i] : 0u) | b;
    f0[i] = b;
    f1[i] = 0u;
    if (f2) f2[i] = 0u;
  }
  if (t < nz0) z0[t] = 0ull;
  if (t < nz1) z1[t] = 0ull;
}

static int fill_grid(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(b, 8192));
}

void launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
}

// ---------------------------------------------------------------------------
// PageRank update kernels.  One wave covers 64 consecutive vertices, so the
// next-frontier words are written whole from a ballot (no atomics).  The
// number of activated vertices is summed per block and added to aux[0] once
// per block: an atomic per wave on that one address serialises (a million
// of them took 10 ms per iteration at n = 2^26).

// Adds each wave's count (held by lane 0) to *aux, one atomic per block.
__device__ __forceinline__ void block_count_add(unsigned long long c, unsigned long long *aux) {
  __shared__ unsigned long long s_c[256 / 64];
  if (lane_id() == 0) s_c[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; w++) t += s_c[w];
    if (t) atomicAdd(aux, t);
  }
}

// Push (pr_kernels.cu:846-860): residual above tol is applied and re-pushed.
__global__ void __launch_bounds__(256)
k_pr_push_update(uint64_t n, float alpha, float tol, const uint32_t *__restrict__ deg,
                 float *rank, float *delta, float *res, uint32_t *next,
                 unsigned long long *aux) {
  const uint64_t words64 = (n + 63) / 64;
  unsigned long long cnt = 0;
  for (uint64_t base = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) & ~63ull;
       base < words64 * 64; base += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = base + lane_id();
    bool act = false;
    if (v < n) {
      const float r = res[v];
      if (r > tol) {
        const uint32_t d = deg[v];
        rank[v] += r;
        delta[v] = d ? r * alpha / (float)d : 0.0f;
        res[v] = 0.0f;
        act = true;
      }
    }
    const uint64_t b = __ballot(act);
    if (lane_id() == 0) {
      next[base / 32] = (uint32_t)b;
      next[base / 32 + 1] = (uint32_t)(b >> 32);
      cnt += __popcll(b);
    }
  }
  block_count_add(cnt, aux);
}

// Pull (pr_kernels.cu:449-476): active vertices take their new sum and stay
// active while they moved by more than tol.
__global__ void __launch_bounds__(256)
k_pr_pull_update(uint64_t n, float alpha, float tol, const uint32_t *__restrict__ outdeg,
                 const uint32_t *__restrict__ cur, float *rank, float *contrib, float *sum,
                 uint32_t *next, unsigned long long *aux) {
  const uint64_t words64 = (n + 63) / 64;
  unsigned long long cnt = 0;
  for (uint64_t base = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) & ~63ull;
       base < words64 * 64; base += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = base + lane_id();
    bool act = false;
    if (v < n && ((cur[v >> 5] >> (v & 31)) & 1u)) {
      const float nr = (1.0f - alpha) + alpha * sum[v];
      const float d = fabsf(nr - rank[v]);
      rank[v] = nr;
      const uint32_t od = outdeg[v];
      contrib[v] = od ? nr / (float)od : 0.0f;
      sum[v] = 0.0f;
      act = d > tol;
    }
    const uint64_t b = __ballot(act);
    if (lane_id() == 0) {
      next[base / 32] = (uint32_t)b;
      next[base / 32 + 1] = (uint32_t)(b >> 32);
      cnt += __popcll(b);
    }
  }
  block_count_add(cnt, aux);
}

__global__ void k_pr_push_init(uint64_t n, float alpha, const uint32_t *__restrict__ deg,
                               float *rank, float *delta, float *res) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = deg[v];
    rank[v] = 1.0f - alpha;
    delta[v] = d ? (1.0f - alpha) * alpha / (float)d : 0.0f;
    res[v] = 0.0f;
  }
}

__global__ void k_pr_pull_init(uint64_t n, const uint32_t *__restrict__ outdeg, float *rank,
                               float *contrib, float *sum) {
  const float r0 = 1.0f / (float)n;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t od = outdeg[v];
    rank[v] = r0;
    contrib[v] = od ? r0 / (float)od : 0.0f;
    sum[v] = 0.0f;
  }
}

// ---------------------------------------------------------------------------

static void stats_begin(thv_stats *st) {
  if (st) memset(st, 0, sizeof *st);
}

static void stats_level(Graph &g, thv_stats *st, const LevelCtr &h) {
  if (!st) return;
  const uint64_t rec = (uint64_t)g.rec;
  const uint64_t words = g.n / 32 + 2;
  const uint64_t e = h.qw[0] + h.qw[1];
  st->edges_traversed += e;
  st->edges_examined += e;
  st->host_bytes += h.qw[0] * rec;
  // device-resident records, one 4-byte state access per edge, 12 bytes
  // (degree, first-record address half, own value) per active vertex, and
  // the frontier bitmap read + cleared once per level.
  st->hbm_bytes += h.qw[1] * rec + 4 * e + 12 * (h.qn[0] + h.qn[1]) + 8 * words;
  st->topdown_host_bytes += h.qw[0] * rec;
  st->topdown_hbm_bytes += h.qw[1] * rec + 4 * e + 12 * (h.qn[0] + h.qn[1]) + 8 * words;
  const uint32_t host_parts = g.store.P - g.store.Pdev;
  st->host_partitions_touched += h.touched[0];
  st->host_partitions_skipped += host_parts - std::min(host_parts, h.touched[0]);
  st->iterations++;
}

static ExpandArgs base_args(Graph &g, Workspace &w) {
  ExpandArgs a = expand_args(g, w.qs);
  a.val = w.val.as<uint32_t>();
  a.vis = w.vis.as<uint32_t>();
  return a;
}
static ExpandArgs base_args(Graph &g) { return base_args(g, g.ws); }

// Level-synchronous loop shared by BFS, CC and SSSP.  front[0] holds the
// first frontier on entry.  Without events (w.ev null) no kernel time is
// taken.
static void value_loop(Graph &g, Workspace &w, Algo algo, thv_stats *st) {
  hipStream_t s = g.stream;
  const uint64_t words = g.n / 32 + 2;
  const bool timed = w.ev[0] && w.ev[1] && g.kernel_timing;
  ExpandArgs a = base_args(g, w);
  LevelCtr *d_ctr = w.qs.ctr.as<LevelCtr>();
  int cur = 0;
  double expand_ms = 0;
  uint32_t launches = 0;
  const uint64_t t0 = now_ns();
  for (uint32_t L = 0;; L++) {
    LevelCtr *c = d_ctr + (L & 1);
    // BFS: the level-L frontier holds the vertices at value L + 1
    launch_compact(g, w.qs, w.front[cur].as<uint32_t>(), c, s, 0,
                   algo == A_BFS ? w.val.as<uint32_t>() : nullptr, L + 1);
    HIPC(hipMemsetAsync(w.front[cur ^ 1].p, 0, words * 4, s));
    a.ctr = c;
    a.next = w.front[cur ^ 1].as<uint32_t>();
    a.level_value = L + 2;
    if (timed) HIPC(hipEventRecord(w.ev[0], s));
    launch_expand(g, algo, a, s);
    if (timed) HIPC(hipEventRecord(w.ev[1], s));
    HIPC(hipMemcpyAsync(&w.h_ctr[L & 1], c, sizeof(LevelCtr), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    const LevelCtr &h = w.h_ctr[L & 1];
    // every launch is counted, the last one (empty frontier, exits at once)
    // included, so launches and time match a kernel trace of the same run
    float ms = 0;
    if (timed) HIPC(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
    expand_ms += ms;
    launches++;
    if (h.qn[0] + h.qn[1] == 0) break;
    stats_level(g, st, h);
    cur ^= 1;
  }
  if (st) {
    st->ms = (now_ns() - t0) * 1e-6;
    st->expand_ms = expand_ms;
    st->expand_launches = launches;
  }
}

static void finish_values(Graph &g, uint32_t *out, thv_stats *st, uint32_t sentinel) {
  if (out && g.n)
    HIPC(hipMemcpy(out, g.ws.val.p, g.n * 4, hipMemcpyDeviceToHost));
  if (st && out) {
    uint64_t r = 0;
    for (uint64_t v = 0; v < g.n; v++) r += out[v] != sentinel;
    st->reached = r;
  }
}

static void check_src(Graph &g, uint32_t src) {
  if (src >= g.n)
    fail(THV_EINVAL, "source vertex " + std::to_string(src) + " out of range (n = " +
                         std::to_string(g.n) + ")");
}

// Direction-optimising BFS on a symmetric graph (Beamer et al., SC'12).
// Each level knows its frontier's size (vertices n_f, edges m_f, per
// residency): level 0 counts it, every later one reads the sums the last
// level's kernels made while placing the vertices.  All edges in HBM:
// with m_u the edges of still-unvisited vertices, the level runs bottom-up
// when m_f > m_u / alpha and returns to top-down once n_f < n / beta and
// the frontier shrinks; once a probe has left candidates, a level also
// runs bottom-up when a probe over them costs less than the top-down
// level's compaction.  Split store: what matters is the host link, so when
// the frontier's host edges are worth it the probe runs first (parent
// hints: every unvisited vertex whose hint is in the frontier joins the
// next level, at level L+1 whichever way the level runs; the rest of the
// unvisited set is counted and sorted into short and long lists), and the
// cheaper direction in link time is taken (below).
//
// A top-down level is a compaction (a full one: no count pass comes
// before it) and an expansion that also writes the discovered vertices'
// level and sums their degrees per residency (A_BFS_NEXT).  A bottom-up
// level is the probe (a sweep of every vertex the first time, the previous
// probe's candidates after that), the full scan of the short lists
// (degree <= D, do_pull_degree) and the early-exit kernel over the long
// ones, which stops at the first frontier neighbour.  edges_traversed keeps
// the algorithm-independent count (sum of degrees of reached vertices);
// edges_examined counts the records actually read.
//
// Host round trips: one stream sync per level (two when a probe's counts
// decide the direction), and no memset or readback copy.  The kernel that
// finishes a level's counters writes them to the host-mapped readback (the
// count pass's scan at level 0, the probe's last launch, a one-wave launch
// after the early-exit kernel or the expansion); the count pass, the probe
// or the compaction clears the next level's frontier bitmap, one of three
// in rotation.
constexpr int AUX_PC = 4;      // aux[4, 10): probe counters
constexpr int HAUX_PC = 8;     // h_aux[8, 14): published probe counters

// Elapsed time between two recorded events that the stream has passed
// (the runtime may not have marked them complete yet).
static float event_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  hipError_t e = hipEventElapsedTime(&ms, a, b);
  if (e == hipErrorNotReady) {
    HIPC(hipEventSynchronize(b));
    e = hipEventElapsedTime(&ms, a, b);
  }
  HIPC(e);
  return ms;
}

static void bfs_direction_optimizing(Graph &g, uint32_t src, thv_stats *st) {
  Workspace &w = g.ws;
  hipStream_t s = g.stream;
  const uint64_t words = bitmap_words(g.n);
  ExpandArgs a = base_args(g);
  LevelCtr *d_ctr = w.qs.ctr.as<LevelCtr>();
  LevelCtr *c_long = d_ctr + 2, *c_short = d_ctr + 3;
  const uint64_t rec = (uint64_t)g.rec;
  const uint32_t D = g.do_pull_degree;
  int cur = 0;  // this level's frontier: front[cur]; next: front[(cur + 1) % 3]
  bool bu = false;
  const bool timed = g.kernel_timing;
  const bool trace = getenv("THV_TRACE") != nullptr;
  const bool host_share = g.store.host_vertices > 0;
  const double link = g.do_link_gbps * 1e9;
  const bool hints = g.store.d_hint.p != nullptr;
  // probe counters: hint hits, unvisited, unvisited with a hint, words
  // written, candidates written, candidates read
  unsigned long long *d_pc = w.aux.as<unsigned long long>() + AUX_PC;
  // Cost terms of the split-store rule, measured on MI355X at R-MAT 26:
  // a short host list scanned in full costs its bytes plus SHORT_NS (a
  // partly shared request), a long one do_vertex_ns (one early-exit read);
  // the HBM side of a top-down level costs TD_NS_ENTRY per queued vertex
  // (compaction write) and TD_NS_EDGE per edge (expansion); a probe sweep
  // PROBE_NS_VERTEX per vertex plus PROBE_FIXED (launches and readback).
  constexpr double SHORT_NS = 0.6, TD_NS_ENTRY = 0.053, TD_NS_EDGE = 0.019,
                   PROBE_NS_VERTEX = 0.0045, PROBE_FIXED = 30e-6;
  // A top-down level also pays its compaction's passes over every tile
  // (TD_FIXED plus TD_NS_TILE_ALL per tile) and TD_NS_TILE per tile that
  // holds frontier vertices; a later probe (candidate list) costs
  // LIST_FIXED for its launches plus LIST_NS_CAND per candidate.  Also
  // measured at R-MAT 26: a 78-vertex top-down level 45 us, a 25K-vertex
  // one 170 us, a list probe level over 25K candidates 60 us.
  constexpr double TD_FIXED = 25e-6, TD_NS_TILE_ALL = 1.2, TD_NS_TILE = 9.7,
                   LIST_FIXED = 60e-6, LIST_NS_CAND = 0.03;
  const double ntiles_d = (double)((g.n + TILE - 1) / TILE);
  uint64_t last_cand = 0;  // candidates the latest probe left
  // HBM side of a top-down level: tiles with frontier vertices taken as nf
  // vertices over the tiles at random
  auto td_hbm_cost = [&](uint64_t nf, uint64_t mf) {
    const double tiles_work = ntiles_d * -std::expm1(-(double)nf / ntiles_d);
    return TD_FIXED + ((double)nf * TD_NS_ENTRY + (double)mf * TD_NS_EDGE +
                       ntiles_d * TD_NS_TILE_ALL + tiles_work * TD_NS_TILE) * 1e-9;
  };
  auto list_cost = [&]() { return (double)last_cand * LIST_NS_CAND * 1e-9 + LIST_FIXED; };
  uint64_t visited_edges = 0, prev_nf = 0;
  double expand_ms = 0, bottom_up_ms = 0, probe_ms = 0, last_probe_ms = 0;
  uint32_t launches = 0, probes = 0;
  uint64_t probe_bytes = 0;  // the parts that do not come from d_pc
  // every level's kernels write the values of the vertices they place and
  // sum the next frontier (c_next): a bottom-up level's probe, short-list
  // and early-exit kernels, a top-down level's A_BFS_NEXT expansion; so only
  // level 0 counts its frontier (fused_next: the next level's sums are in)
  LevelCtr *c_next = d_ctr + 4;
  bool fused_next = false, probed_last = false;
  // a top-down level's expansion, then its counters to the readback: slots
  // 0-4, this level's totals with the partitions its write pass touched and
  // the next frontier's sums (c_next, then reset)
  // (one: the frontier is {src}, expanded from its list without a queue;
  // the launch also writes c's totals and src's value)
  auto expand_next = [&](LevelCtr *c, uint32_t level_value, bool one) {
    a.ctr = c;
    a.level_value = level_value;
    a.deg = g.d_deg.as<uint32_t>();
    a.rbit = g.store.d_rbit.as<uint32_t>();
    a.nctr = c_next;
    if (timed) HIPC(hipEventRecord(w.ev[0], s));
    if (one) launch_expand_one(g, a, src, c, level_value - 1, s);
    else launch_expand(g, A_BFS_NEXT, a, s);
    if (timed) HIPC(hipEventRecord(w.ev[1], s));
    LevelSide ts;
    ts.pub_src = d_ctr;
    ts.pub = w.h_ctr_dev;
    ts.npub = 5;
    ts.zctr = c_next;
    launch_publish(ts, s);
  };
  // the first probe sweeps every vertex; it and every later one write the
  // vertices still unvisited after their hint test (candidates), and a
  // later probe reads only those (cand: the buffer holding the last list)
  int cand = -1;
  const uint64_t ntiles = (g.n + TILE - 1) / TILE;
  // the probe (c_next starts at zero: the count pass or the last early-exit
  // kernel reset it) publishes the long, short and next-frontier counters
  // (slots 2-4) and the probe counters; clear: the bitmap to clear, when no
  // count pass of this level did
  // fsum: the frontier summary for the short-list scan, or null
  auto probe = [&](const uint32_t *front, const uint32_t *hint, uint32_t *next,
                   uint32_t level_value, uint32_t *clear, const uint32_t *fsum) {
    const int out = cand < 0 ? 0 : cand ^ 1;
    QueueSet &q = w.qs;
    LevelSide ps;
    ps.zero = clear;
    ps.nzero = clear ? words : 0;
    ps.pub_src = c_long;
    ps.pub = w.h_ctr_dev + 2;
    ps.npub = 3;
    ps.xsrc = d_pc;
    ps.xpub = w.h_aux_dev + HAUX_PC;
    ps.nx = 6;
    launch_bu_probe(g, q, front, w.vis.as<uint32_t>(), next, hint, D, c_short, c_long, d_pc, s,
                    timed ? w.ev[2] : nullptr, timed ? w.ev[3] : nullptr, c_next, w.val.as<uint32_t>(), level_value,
                    cand < 0 ? nullptr : q.clist[cand].as<uint32_t>(),
                    cand < 0 ? nullptr : q.ccnt[cand].as<uint32_t>(), q.clist[out].as<uint32_t>(),
                    q.ccnt[out].as<uint32_t>(), &ps, fsum);
    // a sweep reads the visited and residency words and writes the tile
    // tables of both sets, a list probe reads its regions' candidate
    // counts (the candidates themselves are counted on the device); both
    // write the new candidate counts
    probe_bytes += (cand < 0 ? 2 * 4 * words + 2 * 24 * ntiles : 4 * 4 * ntiles) + 4 * 4 * ntiles;
    probes++;
    cand = out;
  };
  auto probe_done = [&]() {  // after the stream has drained past the probe
    const float pm = timed ? event_ms(w.ev[2], w.ev[3]) : 0.0f;
    probe_ms += pm;
    last_probe_ms = pm;
    // short lists: an address each, their HBM records, and a frontier word
    // per record (records read from host memory count as host bytes)
    const LevelCtr &ps = w.h_ctr[3];
    const uint64_t sn = ps.qn[0] + ps.qn[1], sr = ps.qw[0] + ps.qw[1];
    probe_bytes += 8 * sn + rec * ps.qw[1] + 4 * sr;
  };
  // Level bookkeeping, run after the next level's count has synchronised
  // the stream (one host round trip per level instead of two).
  struct LevelRec {
    // sure: bottom-up without a readback of the probe's counts (see below);
    // later: level 0 read no counts, h comes from the readback at its end
    // (and its queue was written without a pass over the bitmap); sparse: a
    // top-down level whose queue came from one pass listing its vertices
    bool valid = false, bu = false, probed = false, hinted = false, sure = false, later = false,
         sparse = false;
    uint32_t L = 0;
    uint64_t nf = 0, mf = 0;
    LevelCtr h;
    double td_cost = 0, bu_cost = 0;
  } pend;
  auto short_long_cost = [&](const LevelCtr &pl, const LevelCtr &ps) {
    return (double)ps.qw[0] * rec / link + (double)ps.qn[0] * SHORT_NS * 1e-9 +
           (double)pl.qn[0] * g.do_vertex_ns * 1e-9;
  };
  // An upper bound of any level's bottom-up cost: every host list of 1..16
  // records read in full, and every host list priced as a long one (a
  // level's short host lists, full-scan degree <= 16, and long host lists
  // are subsets of these)
  const double bu_bound = (double)g.store.host_short_edges * rec / link +
                          (double)g.store.host_short_lists * SHORT_NS * 1e-9 +
                          (double)g.store.host_vertices * g.do_vertex_ns * 1e-9;
  // a bound of this level's bottom-up cost: the last level's when it was
  // probed and ran bottom-up (the unvisited set only shrinks), else bu_bound
  double prev_bu_cost = bu_bound;
  auto finish_level = [&](LevelRec &p) {
    if (!p.valid) return;
    if (p.later) p.h = w.h_ctr[p.L & 1];  // published after the expansion
    if (p.bu && (!p.probed || p.sure)) probe_done();
    if (p.sure) p.bu_cost = short_long_cost(w.h_ctr[2], w.h_ctr[3]);
    prev_bu_cost = p.bu && p.probed ? std::min(p.bu_cost, bu_bound) : bu_bound;
    const float ms = timed ? event_ms(w.ev[0], w.ev[1]) : 0.0f;
    LevelCtr hs, hb;  // bottom-up: full-scan and early-exit sets
    memset(&hs, 0, sizeof hs);
    memset(&hb, 0, sizeof hb);
    if (p.bu) {
      hb = w.h_ctr[2];
      if (D > 0) hs = w.h_ctr[3];
      bottom_up_ms += ms;
    } else {
      expand_ms += ms;
      launches++;
    }
    if (trace) {
      fprintf(stderr,
              "[thv] bfs L=%u %s nf=%llu mf=%llu front_host_edges=%llu est_td_ms=%.3f "
              "est_bu_ms=%.3f hint_hits_total=%llu short=%llu short_host_edges=%llu long=%llu "
              "long_host=%llu read_host=%llu read_hbm=%llu kernel_ms=%.3f probe_ms=%.3f\n",
              p.L, p.bu ? "bottom-up" : "top-down", (unsigned long long)p.nf,
              (unsigned long long)p.mf, p.h.qw[0], p.td_cost * 1e3, p.bu_cost * 1e3,
              p.hinted ? w.h_aux[HAUX_PC] : 0ull,
              hs.qn[0] + hs.qn[1], hs.qw[0], hb.qn[0] + hb.qn[1], hb.qn[0],
              p.bu ? hs.qw[0] + hb.rd[0] : p.h.qw[0], p.bu ? hs.qw[1] + hb.rd[1] : p.h.qw[1], ms,
              p.bu || p.probed ? last_probe_ms : 0.0);
    }
    if (st) {
      st->edges_traversed += p.mf;
      st->iterations++;
      st->topdown_host_bytes += p.h.qw[0] * rec;
      st->topdown_hbm_bytes += p.h.qw[1] * rec + 4 * p.mf + 12 * p.nf + 8 * words;
      if (!p.bu) {
        st->edges_examined += p.mf;
        st->host_bytes += p.h.qw[0] * rec;
        st->hbm_bytes += p.h.qw[1] * rec + 4 * p.mf + 12 * p.nf +
                         (p.later ? 0 : p.sparse ? 4 * words : 8 * words);
        // k_touched added them to this level's counters after their count;
        // the publish launch after the expansion copied them out
        st->host_partitions_touched += w.h_ctr[p.L & 1].touched[0];
      } else {
        const uint64_t rh = hs.qw[0] + hb.rd[0], rd = hs.qw[1] + hb.rd[1];
        st->edges_examined += rh + rd;
        st->host_bytes += rh * rec;
        st->bottom_up_host_bytes += rh * rec;
        // early-exit part (the probe's bytes are added at the end): records
        // in HBM, a frontier word per record, 12 bytes per long unvisited
        // list, and the visited / frontier bitmaps
        st->hbm_bytes += hb.rd[1] * rec + 4 * (hb.rd[0] + hb.rd[1]) +
                         12 * (hb.qn[0] + hb.qn[1]) + 12 * words;
        st->host_partitions_touched += hb.touched[0] + hs.touched[0];
        st->bottom_up_levels++;
      }
    }
  };
  const uint64_t t0 = now_ns();
  for (uint32_t L = 0;; L++) {
    LevelCtr *c = d_ctr + (L & 1);
    uint32_t *front = w.front[cur].as<uint32_t>();
    uint32_t *next = w.front[(cur + 1) % 3].as<uint32_t>();
    uint32_t *dead = w.front[(cur + 2) % 3].as<uint32_t>();  // the last level's frontier
    // fused (every level but 0): the last level wrote this frontier's values
    // and summed it (published in slot 4); level 0 counts it (and writes its
    // values) now.  The count pass publishes slots 0-1 (this level's totals
    // and the partitions the last level's write pass touched), resets c_next
    // and clears `dead` for the next level; a fused level has its probe or
    // compaction clear it.
    const bool fused = fused_next;
    // Level 0 from a source whose whole list would not pay for a probe (or,
    // all edges in HBM, would not pass the edge-ratio rule) runs top-down
    // whatever its counts: its list expanded in one launch, no queue, no
    // readback.
    const uint64_t d0 = L == 0 ? g.h_off[src + 1] - g.h_off[src] : 0;
    const bool quick =
        L == 0 && d0 > 0 &&
        (host_share ? (double)d0 * rec / link + td_hbm_cost(1, d0) <=
                          (double)g.n * PROBE_NS_VERTEX * 1e-9 + PROBE_FIXED
                    : (double)d0 <= (double)(g.m - std::min<uint64_t>(g.m, d0)) / g.do_alpha);
    if (!quick && !fused) {
      LevelSide cs;
      cs.zero = dead;
      cs.nzero = words;
      cs.pub_src = d_ctr;
      cs.pub = w.h_ctr_dev;
      cs.npub = 2;
      cs.zctr = c_next;
      launch_compact(g, w.qs, front, c, s, CMP_COUNT_ONLY, w.val.as<uint32_t>(), L + 1, 1,
                     0xFFFFFFFFu, 0, nullptr, 0, nullptr, nullptr, &cs);
    }
    if (quick) {
      // the frontier is {src}: its list expanded directly (the start cleared
      // `dead` and c_next); the publish launch after it copies the counters
      a.next = next;
      expand_next(c, L + 2, true);
      visited_edges += d0;
      prev_nf = 1;
      pend = LevelRec();
      pend.valid = pend.later = true;
      pend.L = L;
      pend.nf = 1;
      pend.mf = d0;
      fused_next = true;
      probed_last = false;
      cur = (cur + 1) % 3;
      continue;
    }
    HIPC(hipStreamSynchronize(s));
    finish_level(pend);
    pend.valid = false;
    const LevelCtr h = fused ? w.h_ctr[4] : w.h_ctr[L & 1];
    uint32_t *clear = fused ? dead : nullptr;
    const uint64_t nf = h.qn[0] + h.qn[1], mf = h.qw[0] + h.qw[1];
    if (fused && probed_last) last_cand = h.cand;
    if (nf == 0) break;
    visited_edges += mf;
    const uint64_t mu = g.m - std::min<uint64_t>(g.m, visited_edges);
    // probed: the bottom-up probe ran (tile set 0 now holds the short
    // unvisited lists, so a top-down level needs a fresh frontier count)
    bool probed = false, hinted = false, sure = false, sparse = false;
    double td_cost = 0, bu_cost = 0;
    const uint32_t *hint = hints ? g.store.d_hint.as<uint32_t>() : nullptr;
    // a frontier of under n / 1024 vertices (THV_OPT_DO_SUMMARY) leaves most
    // summary bits clear: the bottom-up scans then test the summary before a
    // frontier word -- at the traversal's first probe, which scans every
    // unvisited vertex (a later probe's candidates are few by then)
    const bool sparse_front = g.do_summary_ratio > 0 &&
                              (double)nf * g.do_summary_ratio < (double)g.n &&
                              (cand < 0 || g.do_summary_ratio < 1);
    const uint32_t *fsum = nullptr;
    auto summary = [&]() {
      if (sparse_front && !fsum) {
        launch_fsum(g, front, w.fsum.as<uint32_t>(), s);
        fsum = w.fsum.as<uint32_t>();
      }
    };
    if (host_share) {
      // top-down streams the frontier's host records over the link and
      // writes and expands its queue in HBM; the hybrid bottom-up level
      // streams the short host lists in full and pays one small zero-copy
      // read per long unvisited host list.  The probe (hints, short lists,
      // counts of the unvisited set) runs only when top-down costs more
      // than the probe sweep itself.
      td_cost = (double)h.qw[0] * rec / link + td_hbm_cost(nf, mf);
      const double probe_cost =
          cand < 0 ? (double)g.n * PROBE_NS_VERTEX * 1e-9 + PROBE_FIXED : list_cost();
      bu = false;
      if (td_cost > probe_cost) {
        summary();
        probe(front, hint, next, L + 2, clear, fsum);
        clear = nullptr;
        probed = true;
        hinted = hints;
        // the short and long unvisited host lists of this level are subsets
        // of the last level's (the unvisited set only shrinks, degrees are
        // fixed), so its bottom-up cost is at most the last level's, and at
        // most bu_bound: if that is below this level's top-down cost,
        // bottom-up wins without the readback (the counts are read with the
        // level's bookkeeping).  The choice is the one the counts would give.
        if (prev_bu_cost < td_cost) {
          bu = sure = true;
        } else {
          HIPC(hipStreamSynchronize(s));  // the probe published slots 2-4
          probe_done();
          bu_cost = short_long_cost(w.h_ctr[2], w.h_ctr[3]);
          bu = bu_cost < td_cost;
          last_cand = w.h_ctr[4].cand;
        }
      }
    } else {
      // everything in HBM: the edge-ratio rule of Beamer et al.
      if (!bu && (double)mf > (double)mu / g.do_alpha) bu = true;
      else if (bu && (double)nf < (double)g.n / g.do_beta && nf < prev_nf) bu = false;
      // once a probe has left candidates, a bottom-up level costs a list
      // probe over them: taken when cheaper than the top-down compaction
      if (!bu && cand >= 0) {
        td_cost = td_hbm_cost(nf, mf);
        bu_cost = list_cost();
        bu = bu_cost < td_cost;
      }
    }
    prev_nf = nf;
    a.next = next;
    if (!bu) {
      // without a count of this frontier in the tile tables (the probe
      // replaced them, or the count was fused away): a full compaction
      LevelSide zs;
      zs.zero = clear;
      zs.nzero = clear ? words : 0;
      if (nf <= SPARSE_MAX) {
        // a small frontier: listed by one pass over the bitmap, no tile tables
        launch_queue_sparse(g, w.qs, front, c, s, &zs);
        sparse = true;
      } else {
        launch_compact(g, w.qs, front, c, s, probed || fused ? 0 : CMP_WRITE_ONLY, nullptr, 0,
                       1, 0xFFFFFFFFu, 0, nullptr, 0, nullptr, nullptr, &zs);
      }
      expand_next(c, L + 2, false);
    } else {
      if (!probed) {
        summary();
        probe(front, hint, next, L + 2, clear, fsum);
      }
      if (timed) HIPC(hipEventRecord(w.ev[0], s));
      // the probe scanned the short lists already (and, with hints, settled
      // every degree-1 vertex by its hint) and wrote the long ones, per wave
      // region, for the early-exit kernel
      BottomUpArgs b;
      b.llist = w.qs.llist.as<uint32_t>();
      b.loff = w.qs.loff.as<uint32_t>();
      b.nregions = (uint32_t)(div_up(g.n, TILE) * (TILE / PROBE_REGION));
      b.rbit = g.store.d_rbit.as<uint32_t>();
      b.eptr = g.store.d_eptr.as<uint64_t>();
      b.deg = g.d_deg.as<uint32_t>();
      b.front = front;
      b.vis = w.vis.as<uint32_t>();
      b.next = next;
      b.val = w.val.as<uint32_t>();
      b.rd_ctr = c_long;
      b.level_value = L + 2;
      b.nctr = c_next;
      b.fsum = fsum;
      // published after it: slots 2-4 (records read, the short lists'
      // counts, the next frontier's sums); c_next is reset
      b.side.pub_src = c_long;
      b.side.pub = w.h_ctr_dev + 2;
      b.side.npub = 3;
      b.side.zctr = c_next;
      launch_bottom_up(g, b, s);
      if (timed) HIPC(hipEventRecord(w.ev[1], s));
    }
    pend = LevelRec();  // `later` belongs to the quick level 0 only
    pend.valid = true;
    pend.bu = bu;
    pend.probed = probed;
    pend.hinted = hinted;
    pend.sure = sure;
    pend.sparse = sparse;
    pend.L = L;
    pend.nf = nf;
    pend.mf = mf;
    pend.h = h;
    pend.td_cost = td_cost;
    pend.bu_cost = bu_cost;
    fused_next = true;
    probed_last = probed;
    cur = (cur + 1) % 3;
  }
  if (st) {
    st->ms = (now_ns() - t0) * 1e-6;
    st->expand_ms = expand_ms;
    st->expand_launches = launches;
    st->bottom_up_ms = bottom_up_ms;
    if (probes) {
      const unsigned long long *pc = w.h_aux + HAUX_PC;  // published by the last probe
      st->hint_hits = pc[0];
      st->probe_ms = probe_ms;
      st->probe_launches = probes;
      // per unvisited vertex: its hint word (when hints are on) and, when
      // it has one, the frontier word it names, and its degree; the value
      // of each vertex a hint placed; 12 bytes per bitmap word updated
      // (visited written, next read and written); 4 bytes per candidate
      // written, 12 per candidate read (its id, visited and residency words)
      st->probe_hbm_bytes = probe_bytes + (hints ? 4 * pc[1] + 4 * pc[2] : 0) + 4 * pc[1] +
                            4 * pc[0] + 12 * pc[3] + 4 * pc[4] + 12 * pc[5];
      st->hbm_bytes += st->probe_hbm_bytes;
    }
  }
}

void run_bfs(Graph &g, uint32_t src, uint32_t *out, thv_stats *st) {
  check_src(g, src);
  workspace_ensure(g);
  Workspace &w = g.ws;
  hipStream_t s = g.stream;
  const uint64_t words = g.n / 32 + 2;
  stats_begin(st);
  // on a symmetric graph a degree-0 vertex has no in-edges either: it
  // starts visited (its level stays n+1), so bottom-up levels never look at
  // it.  The direction-optimising loop also starts from zero counter slots
  // (the next-frontier sums in slot 4, the sparse-queue list counts) and
  // probe counters (aux[4, 10)).
  const bool dopt = g.sym && g.dir_opt;
  hipLaunchKernelGGL(k_bfs_init, dim3(fill_grid(std::max<uint64_t>(g.n / 4, words))), dim3(256),
                     0, s, w.val.as<uint32_t>(), g.n, src, w.vis.as<uint32_t>(),
                     dopt ? g.d_iso.as<uint32_t>() : nullptr, w.front[0].as<uint32_t>(),
                     w.front[1].as<uint32_t>(), dopt ? w.front[2].as<uint32_t>() : nullptr,
                     words,
                     dopt ? reinterpret_cast<unsigned long long *>(w.qs.ctr.as<LevelCtr>())
                          : nullptr,
                     dopt ? (uint32_t)(CTR_SLOTS * sizeof(LevelCtr) / 8) : 0u,
                     dopt ? w.aux.as<unsigned long long>() + AUX_PC : nullptr,
                     dopt ? 6u : 0u);
  HIPC(hipGetLastError());
  if (dopt) bfs_direction_optimizing(g, src, st);
  else value_loop(g, g.ws, A_BFS, st);
  finish_values(g, out, st, (uint32_t)(g.n + 1));
}

// The reference's top-down BFS (bfs/bfs.cu:55-526: every level expands the
// whole frontier) over the caller's workspace: fills the values and the
// bitmaps with kernels and leaves the levels in w.val.
void bfs_topdown_in(Graph &g, Workspace &w, uint32_t src, thv_stats *st) {
  check_src(g, src);
  hipStream_t s = g.stream;
  const uint64_t words = bitmap_words(g.n);
  stats_begin(st);
  launch_fill_u32(w.val.as<uint32_t>(), (uint32_t)(g.n + 1), g.n, s);
  launch_fill_u32(w.val.as<uint32_t>() + src, 1u, 1, s);
  launch_fill_u32(w.vis.as<uint32_t>(), 0u, words, s);
  launch_fill_u32(w.front[0].as<uint32_t>(), 0u, words, s);
  hipLaunchKernelGGL(k_set_bit, dim3(1), dim3(1), 0, s, w.vis.as<uint32_t>(), src);
  hipLaunchKernelGGL(k_set_bit, dim3(1), dim3(1), 0, s, w.front[0].as<uint32_t>(), src);
  HIPC(hipStreamSynchronize(s));
  value_loop(g, w, A_BFS, st);
}

void run_cc(Graph &g, uint32_t *out, thv_stats *st) {
  workspace_ensure(g);
  Workspace &w = g.ws;
  hipStream_t s = g.stream;
  const uint64_t words = g.n / 32 + 2;
  stats_begin(st);
  if (g.n) {
    hipLaunchKernelGGL(k_iota_u32, dim3(fill_grid(g.n)), dim3(256), 0, s, w.val.as<uint32_t>(),
                       g.n);
  }
  if (g.sym && g.cc_union && g.n) {
    HIPC(hipStreamSynchronize(s));
    if (cc_union_find(g, st)) {
      finish_values(g, out, st, 0xFFFFFFFFu);
      return;
    }
    // a link hit its step guard: start over with label propagation
    stats_begin(st);
    hipLaunchKernelGGL(k_iota_u32, dim3(fill_grid(g.n)), dim3(256), 0, s, w.val.as<uint32_t>(),
                       g.n);
  }
  hipLaunchKernelGGL(k_bits_prefix, dim3(fill_grid(words)), dim3(256), 0, s,
                     w.front[0].as<uint32_t>(), g.n, words);
  HIPC(hipStreamSynchronize(s));
  value_loop(g, g.ws, A_CC, st);
  finish_values(g, out, st, 0xFFFFFFFFu);
}

// SSSP by delta-stepping (Meyer & Sanders 2003) over the same compaction and
// expansion kernels: front[0] is the pending set (vertices whose value
// improved and whose edges were not relaxed since).  Each iteration takes
// the pending vertices with value below the bucket bound T (the write pass
// clears them from the set), relaxes their edges, and the expansion puts
// every improved vertex back.  When no pending vertex is below T, T moves
// up by delta; the loop ends with the set empty.  The fixpoint is the
// frontier Bellman-Ford one (shortest distances are unique); fewer edges
// are relaxed twice because vertices wait for their bucket.
static void sssp_delta(Graph &g, thv_stats *st) {
  Workspace &w = g.ws;
  hipStream_t s = g.stream;
  ExpandArgs a = base_args(g);
  LevelCtr *d_ctr = w.qs.ctr.as<LevelCtr>();
  uint32_t *pend = w.front[0].as<uint32_t>();
  const uint32_t *val = w.val.as<uint32_t>();
  a.next = pend;
  const uint64_t delta = g.sssp_delta;
  uint64_t thr = 1 + delta;  // values are 1 + distance
  double expand_ms = 0;
  uint32_t launches = 0;
  const uint64_t t0 = now_ns();
  for (uint32_t it = 0;; it++) {
    LevelCtr *c = d_ctr + (it & 1);
    const uint32_t t32 = (uint32_t)std::min<uint64_t>(thr, 0xFFFFFFFFull);
    launch_compact(g, w.qs, pend, c, s, 0, nullptr, 0, 1, 0xFFFFFFFFu, 0, val, t32, pend);
    HIPC(hipMemcpyAsync(&w.h_ctr[it & 1], c, sizeof(LevelCtr), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    const LevelCtr h = w.h_ctr[it & 1];
    if (h.qn[0] + h.qn[1] == 0) {
      // bucket done: the smallest pending value opens the next one (none
      // pending: done)
      uint32_t *d_min = reinterpret_cast<uint32_t *>(w.aux.as<unsigned long long>() + 3);
      launch_fill_u32(d_min, 0xFFFFFFFFu, 1, s);
      launch_compact(g, w.qs, pend, c, s, CMP_COUNT_ONLY, nullptr, 0, 1, 0xFFFFFFFFu, 0, val,
                     0xFFFFFFFFu, nullptr, d_min);
      HIPC(hipMemcpyAsync(w.h_aux + 3, d_min, 4, hipMemcpyDeviceToHost, s));
      HIPC(hipStreamSynchronize(s));
      const uint32_t lo = reinterpret_cast<const uint32_t *>(w.h_aux + 3)[0];
      if (lo == 0xFFFFFFFFu) break;
      thr = (uint64_t)lo + delta;
      continue;
    }
    a.ctr = c;
    HIPC(hipEventRecord(w.ev[0], s));
    launch_expand(g, A_SSSP, a, s);
    HIPC(hipEventRecord(w.ev[1], s));
    HIPC(hipStreamSynchronize(s));
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
    expand_ms += ms;
    launches++;
    stats_level(g, st, h);
  }
  if (st) {
    st->ms = (now_ns() - t0) * 1e-6;
    st->expand_ms = expand_ms;
    st->expand_launches = launches;
  }
}

void run_sssp(Graph &g, uint32_t src, uint32_t *out, thv_stats *st) {
  check_src(g, src);
  if (g.rec != 8) fail(THV_ENOTSUP, "SSSP needs a weighted graph (8-byte records)");
  workspace_ensure(g);
  Workspace &w = g.ws;
  hipStream_t s = g.stream;
  const uint64_t words = g.n / 32 + 2;
  stats_begin(st);
  launch_fill_u32(w.val.as<uint32_t>(), (uint32_t)(g.n + 1), g.n, s);
  launch_fill_u32(w.val.as<uint32_t>() + src, 1u, 1, s);
  HIPC(hipMemsetAsync(w.front[0].p, 0, words * 4, s));
  hipLaunchKernelGGL(k_set_bit, dim3(1), dim3(1), 0, s, w.front[0].as<uint32_t>(), src);
  HIPC(hipStreamSynchronize(s));
  if (g.sssp_delta) sssp_delta(g, st);
  else value_loop(g, g.ws, A_SSSP, st);
  finish_values(g, out, st, (uint32_t)(g.n + 1));
}

void run_pr(Graph &g, int mode, float alpha, float tol, int max_iter, float *out,
            thv_stats *st) {
  if (mode != THV_PR_PUSH && mode != THV_PR_PULL) fail(THV_EINVAL, "bad PageRank mode");
  if (mode == THV_PR_PULL && !g.csc)
    fail(THV_ENOTSUP, "PageRank pull needs a CSC graph (in-edges + out-degrees)");
  if (mode == THV_PR_PUSH && g.csc)
    fail(THV_ENOTSUP, "PageRank push needs a CSR graph (out-edges)");
  if (max_iter < 0) fail(THV_EINVAL, "max_iter must be >= 0");
  workspace_ensure(g);
  Workspace &w = g.ws;
  hipStream_t s = g.stream;
  const uint64_t n = g.n;
  const uint64_t words = n / 32 + 2;
  if (w.f0.bytes < (n + 1) * 4) {
    w.f0.alloc((n + 1) * 4);
    w.f1.alloc((n + 1) * 4);
    w.f2.alloc((n + 1) * 4);
  }
  stats_begin(st);
  float *rank = w.f0.as<float>(), *f1 = w.f1.as<float>(), *f2 = w.f2.as<float>();
  const int fg = fill_grid(n);
  if (n) {
    if (mode == THV_PR_PUSH)
      hipLaunchKernelGGL(k_pr_push_init, dim3(fg), dim3(256), 0, s, n, alpha,
                         g.d_deg.as<uint32_t>(), rank, f1, f2);
    else
      hipLaunchKernelGGL(k_pr_pull_init, dim3(fg), dim3(256), 0, s, n,
                         g.d_outdeg.as<uint32_t>(), rank, f1, f2);
  }
  hipLaunchKernelGGL(k_bits_prefix, dim3(fill_grid(words)), dim3(256), 0, s,
                     w.front[0].as<uint32_t>(), n, words);
  HIPC(hipMemsetAsync(w.front[1].p, 0, words * 4, s));
  HIPC(hipStreamSynchronize(s));

  ExpandArgs a = base_args(g);
  a.pr_src = f1;  // push: delta, pull: contribution
  a.pr_dst = f2;  // push: residual, pull: sum
  if (mode == THV_PR_PUSH) {
    ensure_pr_hubs(g);
    a.hubs = g.d_hubs.as<uint32_t>();
    a.hubbit = g.d_hubbit.as<uint32_t>();
    a.nhubs = g.nhubs;
  }
  LevelCtr *d_ctr = w.qs.ctr.as<LevelCtr>();
  unsigned long long *aux = w.aux.as<unsigned long long>();
  const Algo algo = mode == THV_PR_PUSH ? A_PR_PUSH : A_PR_PULL;
  int cur = 0;
  double expand_ms = 0;
  const uint64_t t0 = now_ns();
  const int ug = (int)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 4096));
  for (int it = 0; it < max_iter; it++) {
    LevelCtr *c = d_ctr + (it & 1);
    launch_compact(g, w.qs, w.front[cur].as<uint32_t>(), c, s);
    a.ctr = c;
    HIPC(hipEventRecord(w.ev[0], s));
    launch_expand(g, algo, a, s);
    HIPC(hipEventRecord(w.ev[1], s));
    HIPC(hipMemsetAsync(aux, 0, 8, s));
    if (mode == THV_PR_PUSH)
      hipLaunchKernelGGL(k_pr_push_update, dim3(ug), dim3(256), 0, s, n, alpha, tol,
                         g.d_deg.as<uint32_t>(), rank, f1, f2,
                         w.front[cur ^ 1].as<uint32_t>(), aux);
    else
      hipLaunchKernelGGL(k_pr_pull_update, dim3(ug), dim3(256), 0, s, n, alpha, tol,
                         g.d_outdeg.as<uint32_t>(), w.front[cur].as<uint32_t>(), rank, f1,
                         f2, w.front[cur ^ 1].as<uint32_t>(), aux);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(&w.h_ctr[it & 1], c, sizeof(LevelCtr), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(w.h_aux, aux, 8, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
    expand_ms += ms;
    const LevelCtr &h = w.h_ctr[it & 1];
    stats_level(g, st, h);
    if (st) {
      // push: 4-byte atomic read+write of the residual per edge in place of
      // the state read; pull: 4-byte contribution read per edge.
      const uint64_t e = h.qw[0] + h.qw[1];
      st->hbm_bytes += (mode == THV_PR_PUSH ? 4 : 0) * e;
      st->hbm_bytes += 20 * n;  // update pass over rank / delta / residual
    }
    cur ^= 1;
    if (st) st->expand_launches++;
    if (w.h_aux[0] == 0) break;
  }
  if (st) {
    st->ms = (now_ns() - t0) * 1e-6;
    st->expand_ms = expand_ms;
    st->reached = n;
  }
  if (out && n) HIPC(hipMemcpy(out, rank, n * 4, hipMemcpyDeviceToHost));
}

}  // namespace thv
