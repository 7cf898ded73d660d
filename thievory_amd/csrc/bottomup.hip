// bottomup.hip -- bottom-up BFS level for symmetric graphs.
//
// Direction-optimising BFS (Beamer, Asanovic, Patterson, SC'12): when the
// frontier is large, every unvisited vertex looks for a parent among its
// own neighbours and stops at the first one in the frontier, instead of the
// frontier pushing along all of its edges.  Levels are the same as
// top-down's (the reference only has the top-down rule, bfs_kernels.cu:
// 29-44); what changes is how many edge records cross the host link.
//
// One group of 16 lanes per unvisited vertex (queue from the compaction
// of the inverted visited bitmap): the group reads 16 consecutive records
// (64 bytes of 4-byte records) per step and tests the frontier bitmap; a
// ballot ends the scan at the first hit.  Host and HBM queue entries are
// interleaved like the expansion chunks.  Records read are counted per
// residency for the byte accounting.
#include "internal.hpp"
#include "scan.hpp"

namespace thv {

namespace {

constexpr int BBS = 256;
constexpr int G = 16;  // lanes per vertex

template <int REC>
__global__ void __launch_bounds__(BBS) k_bottom_up(BottomUpArgs a) {
  const uint32_t lane = threadIdx.x & 63, gl = lane & (G - 1), gw = lane / G;
  const unsigned long long group = (blockIdx.x * (unsigned long long)BBS + threadIdx.x) / G;
  const unsigned long long ngroups = (unsigned long long)gridDim.x * BBS / G;
  const unsigned long long nh = a.ctr->qn[Q_HOST], nd = a.ctr->qn[Q_DEV];
  const unsigned long long total = nh + nd;
  unsigned long long rd[2] = {0, 0};
  for (unsigned long long i = group; i < total; i += ngroups) {
    const unsigned long long d0 = i * nd / total, d1 = (i + 1) * nd / total;
    const int q = d1 > d0 ? Q_DEV : Q_HOST;
    const unsigned long long idx = q == Q_DEV ? d0 : i - d0;
    const uint32_t v = a.qv[q][idx];
    const uint32_t d = a.deg[v];
    const unsigned long long base = a.eptr[v];
    bool found = false;
    uint32_t k = 0;
    for (; k < d; k += G) {
      const uint32_t e = k + gl;
      bool hit = false;
      if (e < d) {
        const uint32_t w = *reinterpret_cast<const uint32_t *>(base + (unsigned long long)e * REC);
        hit = (a.front[w >> 5] >> (w & 31)) & 1u;
      }
      const unsigned long long b = __ballot(hit);
      if ((b >> (gw * G)) & 0xFFFFull) {
        found = true;
        break;
      }
    }
    if (gl == 0) {
      rd[q] += found ? min(d, k + G) : d;
      if (found) {
        a.val[v] = a.level_value;
        atomicOr(&a.next[v >> 5], 1u << (v & 31));
        atomicOr(&a.vis[v >> 5], 1u << (v & 31));
      }
    }
  }
  const unsigned long long h = wave_sum(rd[Q_HOST]), dv = wave_sum(rd[Q_DEV]);
  if (lane == 0) {
    if (h) atomicAdd(&a.rd_ctr->rd[Q_HOST], h);
    if (dv) atomicAdd(&a.rd_ctr->rd[Q_DEV], dv);
  }
}

}  // namespace

void launch_bottom_up(Graph &g, const BottomUpArgs &a, hipStream_t s) {
  const int grid = 8 * g.cus;
  if (g.rec == 8) hipLaunchKernelGGL(k_bottom_up<8>, dim3(grid), dim3(BBS), 0, s, a);
  else hipLaunchKernelGGL(k_bottom_up<4>, dim3(grid), dim3(BBS), 0, s, a);
  HIPC(hipGetLastError());
}

}  // namespace thv
