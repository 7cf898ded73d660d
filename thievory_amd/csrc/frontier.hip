// frontier.hip -- per-level frontier compaction into the two work queues.
//
// Replaces the reference's per-level chain setStaticNDemandFrontiers ->
// thrust::reduce x3 -> CalculateActiveEdgesPerPartition ->
// CalculateActiveEdgesRatio -> SplitZeroCopyNFilterFrontiers ->
// thrust::exclusive_scan -> setStaticList / setDemandList
// (src/common.cuh:68-183, src/bfs/bfs.cu:56-153) with three kernels and no
// host round trip, over fixed tiles of 4096 vertices:
//   k_tcount  one block per tile, one wave per 1024-vertex sub-tile: active
//             vertices (degree > 0) and their edges, split by residency
//   k_tscan   one block: exclusive tile bases per queue, totals
//   k_twrite  one block per tile with work: ordered queue entries (vertex
//             id, first edge in the level's work space), the chunk-start
//             table the expansion kernel indexes by chunk, and a count of
//             the partitions touched (stamped, so each counts once)
// A host partition with no active vertex produces no entry, so its edges are
// never read (the per-partition activity filter).  Queues come out sorted by
// vertex id, which makes the zero-copy reads of a dense host frontier
// sequential.  64 vertices whose frontier bits are all clear cost one
// ballot.
#include "internal.hpp"
#include "scan.hpp"

namespace thv {

namespace {

constexpr int CBS = 256;
constexpr uint32_t SUB = TILE / (CBS / 64);  // vertices per wave
constexpr uint32_t ROUNDS = SUB / 64;

struct CompactArgs {
  const uint32_t *front, *rbit, *deg, *vpart;
  uint64_t n;
  uint32_t ntiles;
  uint32_t *tcnt, *tbc;
  unsigned long long *tdeg, *tbe;
  uint32_t *qv[2];
  unsigned long long *qe[2];
  uint32_t *qc[2];
  uint32_t *ptouch;
  uint32_t stamp;
  LevelCtr *ctr;
  uint32_t C;
  uint32_t invert;     // select vertices whose bit is clear
  uint32_t no_chunks;  // skip the chunk-start table
};

// Lane's vertex: active (frontier bit, degree > 0), residency queue, degree.
__device__ __forceinline__ void vertex_state(const CompactArgs &a, uint64_t v, bool &act,
                                             int &q, uint32_t &d) {
  act = false;
  q = Q_HOST;
  d = 0;
  if (v < a.n && (((a.front[v >> 5] >> (v & 31)) & 1u) ^ a.invert)) {
    d = a.deg[v];
    act = d != 0;
    if (act && ((a.rbit[v >> 5] >> (v & 31)) & 1u)) q = Q_DEV;
  }
}

__global__ void __launch_bounds__(CBS) k_tcount(CompactArgs a) {
  __shared__ unsigned long long s[CBS / 64][4];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (uint32_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const uint64_t base = (uint64_t)t * TILE + (uint64_t)wv * SUB;
    uint32_t c[2] = {0, 0};
    unsigned long long e[2] = {0, 0};
    for (uint32_t r = 0; r < ROUNDS; r++) {
      bool act;
      int q;
      uint32_t d;
      vertex_state(a, base + r * 64 + lane, act, q, d);
      if (act) {
        c[q]++;
        e[q] += d;
      }
    }
    const unsigned long long v0 = wave_sum((unsigned long long)c[0]);
    const unsigned long long v1 = wave_sum((unsigned long long)c[1]);
    const unsigned long long v2 = wave_sum(e[0]);
    const unsigned long long v3 = wave_sum(e[1]);
    if (lane == 0) {
      s[wv][0] = v0;
      s[wv][1] = v1;
      s[wv][2] = v2;
      s[wv][3] = v3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      unsigned long long x = 0;
      for (int w = 0; w < CBS / 64; w++) x += s[w][threadIdx.x];
      if (threadIdx.x < 2) a.tcnt[2 * t + threadIdx.x] = (uint32_t)x;
      else a.tdeg[2 * t + threadIdx.x - 2] = x;
    }
    __syncthreads();
  }
}

constexpr int SBS = 1024;

__global__ void __launch_bounds__(SBS) k_tscan(CompactArgs a) {
  __shared__ unsigned long long sh[4][SBS / 64 + 1];
  unsigned long long carry[4] = {0, 0, 0, 0};
  for (uint32_t b = 0; b < a.ntiles; b += SBS) {
    const uint32_t t = b + threadIdx.x;
    unsigned long long x[4] = {0, 0, 0, 0};
    if (t < a.ntiles) {
      x[0] = a.tcnt[2 * t];
      x[1] = a.tcnt[2 * t + 1];
      x[2] = a.tdeg[2 * t];
      x[3] = a.tdeg[2 * t + 1];
    }
    unsigned long long ex[4], tot[4];
#pragma unroll
    for (int i = 0; i < 4; i++) ex[i] = block_excl_scan<SBS>(x[i], sh[i], &tot[i]);
    if (t < a.ntiles) {
      a.tbc[2 * t] = (uint32_t)(carry[0] + ex[0]);
      a.tbc[2 * t + 1] = (uint32_t)(carry[1] + ex[1]);
      a.tbe[2 * t] = carry[2] + ex[2];
      a.tbe[2 * t + 1] = carry[3] + ex[3];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) carry[i] += tot[i];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.ctr->qn[0] = carry[0];
    a.ctr->qn[1] = carry[1];
    a.ctr->qw[0] = carry[2];
    a.ctr->qw[1] = carry[3];
    a.ctr->touched[0] = 0;
    a.ctr->touched[1] = 0;
    a.ctr->rd[0] = 0;
    a.ctr->rd[1] = 0;
  }
}

__global__ void __launch_bounds__(CBS) k_twrite(CompactArgs a) {
  __shared__ unsigned long long s[CBS / 64][4];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long lt = lanemask_lt();
  for (uint32_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    if (a.tcnt[2 * t] == 0 && a.tcnt[2 * t + 1] == 0) continue;
    const uint64_t base = (uint64_t)t * TILE + (uint64_t)wv * SUB;
    // pass 1: this wave's share, to place it after the waves before it
    uint32_t c[2] = {0, 0};
    unsigned long long e[2] = {0, 0};
    for (uint32_t r = 0; r < ROUNDS; r++) {
      bool act;
      int q;
      uint32_t d;
      vertex_state(a, base + r * 64 + lane, act, q, d);
      if (act) {
        c[q]++;
        e[q] += d;
      }
    }
    const unsigned long long w0 = wave_sum((unsigned long long)c[0]);
    const unsigned long long w1 = wave_sum((unsigned long long)c[1]);
    const unsigned long long w2 = wave_sum(e[0]);
    const unsigned long long w3 = wave_sum(e[1]);
    if (lane == 0) {
      s[wv][0] = w0;
      s[wv][1] = w1;
      s[wv][2] = w2;
      s[wv][3] = w3;
    }
    __syncthreads();
    unsigned long long cb[2] = {a.tbc[2 * t], a.tbc[2 * t + 1]};
    unsigned long long eb[2] = {a.tbe[2 * t], a.tbe[2 * t + 1]};
    for (uint32_t w = 0; w < wv; w++) {
      cb[0] += s[w][0];
      cb[1] += s[w][1];
      eb[0] += s[w][2];
      eb[1] += s[w][3];
    }
    // pass 2: ordered entries
    for (uint32_t r = 0; r < ROUNDS; r++) {
      const uint64_t v = base + r * 64 + lane;
      bool act;
      int q;
      uint32_t d;
      vertex_state(a, v, act, q, d);
      const unsigned long long bh = __ballot(act && q == Q_HOST);
      const unsigned long long bd = __ballot(act && q == Q_DEV);
      if (!(bh | bd)) continue;
      const unsigned long long xh = (act && q == Q_HOST) ? d : 0ull;
      const unsigned long long xd = (act && q == Q_DEV) ? d : 0ull;
      const unsigned long long ih = wave_incl_scan(xh), id = wave_incl_scan(xd);
      if (act) {
        const unsigned long long idx =
            cb[q] + __popcll((q == Q_HOST ? bh : bd) & lt);
        const unsigned long long e0 = eb[q] + (q == Q_HOST ? ih - xh : id - xd);
        a.qv[q][idx] = (uint32_t)v;
        a.qe[q][idx] = e0;
        if (!a.no_chunks) {
          uint32_t *qc = a.qc[q];
          for (unsigned long long ch = (e0 + a.C - 1) / a.C; ch * a.C < e0 + d; ch++)
            qc[ch] = (uint32_t)idx;
        }
      }
      // partitions touched: the first lane of every run of equal partition
      // ids stamps it; a stamp that changes counts the partition once
      const uint32_t p = act ? a.vpart[v] : 0xFFFFFFFFu;
      const uint32_t pp = __shfl_up(p, 1, 64);
      if (act && (lane == 0 || pp != p)) {
        if (atomicExch(&a.ptouch[p], a.stamp) != a.stamp) atomicAdd(&a.ctr->touched[q], 1u);
      }
      cb[0] += __popcll(bh);
      cb[1] += __popcll(bd);
      eb[0] += __shfl(ih, 63, 64);
      eb[1] += __shfl(id, 63, 64);
    }
    __syncthreads();
  }
}

}  // namespace

void launch_compact(Graph &g, QueueSet &w, const uint32_t *front, LevelCtr *ctr,
                    hipStream_t s, int mode) {
  Store &st = g.store;
  CompactArgs a;
  memset(&a, 0, sizeof a);
  a.front = front;
  a.rbit = st.d_rbit.as<uint32_t>();
  a.deg = g.d_deg.as<uint32_t>();
  a.vpart = st.d_vpart.as<uint32_t>();
  a.n = g.n;
  a.ntiles = (uint32_t)(g.n / TILE + 1);
  a.tcnt = w.tcnt.as<uint32_t>();
  a.tdeg = w.tdeg.as<unsigned long long>();
  a.tbc = w.tbc.as<uint32_t>();
  a.tbe = w.tbe.as<unsigned long long>();
  for (int q = 0; q < 2; q++) {
    a.qv[q] = w.qv[q].as<uint32_t>();
    a.qe[q] = w.qe[q].as<unsigned long long>();
    a.qc[q] = w.qc[q].as<uint32_t>();
  }
  a.ptouch = w.ptouch.as<uint32_t>();
  a.stamp = ++w.stamp;
  if (a.stamp == 0) {  // wrapped: old stamps could collide
    HIPC(hipMemsetAsync(w.ptouch.p, 0, (uint64_t)(w.P + 1) * 4, s));
    a.stamp = w.stamp = 1;
  }
  a.ctr = ctr;
  a.C = w.C;
  a.invert = (mode & CMP_INVERT) ? 1u : 0u;
  a.no_chunks = (mode & CMP_NO_CHUNKS) ? 1u : 0u;
  const uint32_t grid = a.ntiles < 65535u ? a.ntiles : 65535u;
  // WRITE_ONLY continues a COUNT_ONLY call on the same bitmap and counters:
  // the tile bases and totals are still in place
  if (!(mode & CMP_WRITE_ONLY)) {
    hipLaunchKernelGGL(k_tcount, dim3(grid), dim3(CBS), 0, s, a);
    hipLaunchKernelGGL(k_tscan, dim3(1), dim3(SBS), 0, s, a);
  }
  if (!(mode & CMP_COUNT_ONLY))
    hipLaunchKernelGGL(k_twrite, dim3(grid), dim3(CBS), 0, s, a);
  HIPC(hipGetLastError());
}

ExpandArgs expand_args(const Graph &g, const QueueSet &qs) {
  ExpandArgs a;
  memset(&a, 0, sizeof a);
  for (int q = 0; q < 2; q++) {
    a.qv[q] = qs.qv[q].as<uint32_t>();
    a.qe[q] = qs.qe[q].as<unsigned long long>();
    a.qc[q] = qs.qc[q].as<uint32_t>();
  }
  a.eptr = g.store.d_eptr.as<uint64_t>();
  a.C = qs.C;
  a.rec = g.rec;
  return a;
}

}  // namespace thv
