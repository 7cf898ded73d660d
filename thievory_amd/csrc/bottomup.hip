// bottomup.hip -- bottom-up BFS level for symmetric graphs.
//
// Direction-optimising BFS (Beamer, Asanovic, Patterson, SC'12): when the
// frontier is large, every unvisited vertex looks for a parent among its
// own neighbours and stops at the first one in the frontier, instead of the
// frontier pushing along all of its edges.  Levels are the same as
// top-down's (the reference only has the top-down rule, bfs_kernels.cu:
// 29-44); what changes is how many edge records cross the host link.
//
// One group of G = 8 lanes per unvisited vertex with a long list (the probe
// wrote them per wave region, in vertex order, and k_lbase the regions'
// bases): the group reads 8 consecutive records (32 bytes of 4-byte
// records) per step and tests the frontier bitmap; a ballot ends the scan
// at the first hit.  Every wave takes an equal run of the lists in region
// order, wherever they cluster (on R-MAT graphs the long lists crowd into
// the regions of the low-numbered vertices).  Records read are counted per
// residency for the byte accounting.
#include <algorithm>

#include "internal.hpp"
#include "scan.hpp"
#include "side.hpp"

namespace thv {

namespace {

constexpr int BBS = 256;
// lanes per vertex: 8 measured best at R-MAT 26 (64 random roots 2.2% less
// time than with 16, vertex 1 the same; 4 and 32 slower)
constexpr int G = 8;
constexpr unsigned long long GM = (1ull << G) - 1ull;  // a group's ballot bits

template <int REC, bool SUM>
__global__ void __launch_bounds__(BBS) k_bottom_up(BottomUpArgs a) {
  __shared__ unsigned long long sh[BBS / 64];
  const uint32_t lane = threadIdx.x & 63, gl = lane & (G - 1), gw = lane / G;
  const uint64_t nwaves = (uint64_t)gridDim.x * (BBS / 64);
  const uint64_t wave = (uint64_t)blockIdx.x * (BBS / 64) + (threadIdx.x >> 6);
  unsigned long long rd[2] = {0, 0}, fn[2] = {0, 0}, fe[2] = {0, 0};
  // this wave's run [s, e) of the T lists
  const uint64_t T = a.loff[a.nregions];
  const uint64_t per = (T + nwaves - 1) / nwaves;
  const uint64_t s = wave * per, e = s + per < T ? s + per : T;
  if (s < e) {
    // the region holding list s: the last region whose base is <= s
    // (64 candidates per round; bases are non-decreasing and loff[nregions]
    // = T > s)
    uint32_t lo = 0, hi = a.nregions;
    while (hi - lo > 1) {
      const uint32_t step = (hi - lo + 63) / 64;
      const uint64_t p = (uint64_t)lo + (uint64_t)lane * step;
      const bool le = p < hi && a.loff[p] <= s;
      const uint32_t c = (uint32_t)__popcll(__ballot(le));  // lane 0: p = lo
      lo += (c - 1) * step;
      if ((uint64_t)lo + step < hi) hi = lo + step;
    }
    uint32_t rg = lo;  // per group from here on: the region of its list
    for (uint64_t i0 = s; i0 < e; i0 += 64 / G) {
      const uint64_t i = i0 + gw;
      if (i >= e) break;
      // move on to the region holding list i, G bases per step
      for (;;) {
        const uint64_t r = (uint64_t)rg + 1 + gl;
        const bool le = r <= a.nregions && a.loff[r] <= i;
        const uint32_t c = (uint32_t)__popcll((__ballot(le) >> (gw * G)) & GM);
        rg += c;
        if (c < G) break;
      }
      const uint32_t v = a.llist[(uint64_t)rg * PROBE_REGION + (i - a.loff[rg])];
      const int q = (a.rbit[v >> 5] >> (v & 31)) & 1u ? Q_DEV : Q_HOST;
      const uint32_t d = a.deg[v];
      const unsigned long long base = a.eptr[v];
      bool found = false;
      uint32_t k = 0;
      for (; k < d; k += G) {
        const uint32_t x = k + gl;
        bool hit = false;
        if (x < d) {
          const uint32_t w = *reinterpret_cast<const uint32_t *>(base + (unsigned long long)x * REC);
          // with a frontier summary, its bit first (an L2-resident word)
          const bool maybe = !SUM || ((a.fsum[w >> (FSUM_SHIFT + 5)] >> ((w >> FSUM_SHIFT) & 31)) & 1u);
          hit = maybe && ((a.front[w >> 5] >> (w & 31)) & 1u);
        }
        const unsigned long long b = __ballot(hit);
        if ((b >> (gw * G)) & GM) {
          found = true;
          break;
        }
      }
      if (gl == 0) {
        rd[q] += found ? min(d, k + G) : d;
        if (found) {
          atomicOr(&a.next[v >> 5], 1u << (v & 31));
          atomicOr(&a.vis[v >> 5], 1u << (v & 31));
          if (a.nctr) {  // else the next level's count writes the level
            a.val[v] = a.level_value;
            fn[q] += 1;
            fe[q] += d;
          }
        }
      }
    }
  }
  // one atomic per counter and block; every thread takes part in the sums
  const unsigned long long h = block_sum<BBS>(rd[Q_HOST], sh);
  const unsigned long long dv = block_sum<BBS>(rd[Q_DEV], sh);
  if (threadIdx.x == 0) {
    if (h) atomicAdd(&a.rd_ctr->rd[Q_HOST], h);
    if (dv) atomicAdd(&a.rd_ctr->rd[Q_DEV], dv);
  }
  if (a.nctr) {
    const unsigned long long x0 = block_sum<BBS>(fn[0], sh), x1 = block_sum<BBS>(fn[1], sh);
    const unsigned long long y0 = block_sum<BBS>(fe[0], sh), y1 = block_sum<BBS>(fe[1], sh);
    if (threadIdx.x == 0) {
      if (x0) atomicAdd(&a.nctr->qn[0], x0);
      if (x1) atomicAdd(&a.nctr->qn[1], x1);
      if (y0) atomicAdd(&a.nctr->qw[0], y0);
      if (y1) atomicAdd(&a.nctr->qw[1], y1);
    }
  }
}

// The level's counters to the host-mapped readback once the early-exit
// kernel has finished: one wave, one tiny launch in place of the copies.
// (A last-block count inside k_bottom_up would need a device-scope
// acquire-release per block, an L2 write-back and invalidate each, which
// cost more than this launch.)
__global__ void __launch_bounds__(64) k_publish(LevelSide s) { side_publish(s, threadIdx.x); }

// Parent hints from the keys of the hint-build expansion: the low word of
// a key is ~(neighbour of largest degree, smallest id among ties).
__global__ void __launch_bounds__(BBS) k_hint_finish(const unsigned long long *key, uint64_t n,
                                                    uint32_t *hint) {
  for (uint64_t v = blockIdx.x * (uint64_t)BBS + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * BBS) {
    const unsigned long long k = key[v];
    hint[v] = k ? 0xFFFFFFFFu - (uint32_t)k : NO_HINT;
  }
}

}  // namespace

void launch_publish(const LevelSide &side, hipStream_t s) {
  if (side.zero) fail(THV_ESTATE, "the publish launch clears no words");
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, s, side);
  HIPC(hipGetLastError());
}

void launch_bottom_up(Graph &g, const BottomUpArgs &a, hipStream_t s) {
  if (a.side.zero) fail(THV_ESTATE, "the early-exit kernel clears no words");
  const int grid = g.shape.probe_bpc * g.cus;
  if (g.rec == 8) {
    if (a.fsum) hipLaunchKernelGGL((k_bottom_up<8, true>), dim3(grid), dim3(BBS), 0, s, a);
    else hipLaunchKernelGGL((k_bottom_up<8, false>), dim3(grid), dim3(BBS), 0, s, a);
  } else {
    if (a.fsum) hipLaunchKernelGGL((k_bottom_up<4, true>), dim3(grid), dim3(BBS), 0, s, a);
    else hipLaunchKernelGGL((k_bottom_up<4, false>), dim3(grid), dim3(BBS), 0, s, a);
  }
  if (a.side.pub || a.side.xpub || a.side.zctr)
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, s, a.side);
  HIPC(hipGetLastError());
}

// Parent hints: for every vertex with edges, its neighbour of largest
// degree (ties: smallest id).  BFS reaches high-degree vertices early, so in
// a bottom-up level this one HBM word usually names a frontier neighbour.
// One edge-balanced expansion over every list (HBM and zero-copy host
// records alike), once per store configuration, outside any traversal.
void build_parent_hints(Graph &g) {
  Workspace &w = g.ws;
  workspace_ensure(g);
  hipStream_t s = g.stream;
  DevMem key;
  key.alloc(g.n * 8);
  HIPC(hipMemsetAsync(key.p, 0, g.n * 8, s));
  HIPC(hipMemsetAsync(w.front[0].p, 0, bitmap_words(g.n) * 4, s));
  LevelCtr *c = w.qs.ctr.as<LevelCtr>();
  launch_compact(g, w.qs, w.front[0].as<uint32_t>(), c, s, CMP_INVERT);  // every vertex
  ExpandArgs a = expand_args(g, w.qs);
  a.ctr = c;
  a.deg = g.d_deg.as<uint32_t>();
  a.key64 = key.as<unsigned long long>();
  launch_expand(g, A_HINT, a, s);
  g.store.d_hint.alloc(vertex_pad(g.n) * 4);
  if (vertex_pad(g.n) > g.n)  // the padding: no hint
    HIPC(hipMemsetAsync(g.store.d_hint.as<uint32_t>() + g.n, 0xFF, (vertex_pad(g.n) - g.n) * 4, s));
  const int grid = (int)std::min<uint64_t>(div_up(g.n, BBS), 16ull * g.cus);
  hipLaunchKernelGGL(k_hint_finish, dim3(grid ? grid : 1), dim3(BBS), 0, s,
                     key.as<unsigned long long>(), g.n, g.store.d_hint.as<uint32_t>());
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(s));
}

}  // namespace thv
