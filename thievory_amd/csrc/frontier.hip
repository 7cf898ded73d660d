// frontier.hip -- per-level frontier compaction, partition by partition.
//
// Replaces the reference's per-level chain setStaticNDemandFrontiers ->
// thrust::reduce x3 -> CalculateActiveEdgesPerPartition ->
// CalculateActiveEdgesRatio -> SplitZeroCopyNFilterFrontiers ->
// thrust::exclusive_scan -> setStaticList / setDemandList
// (src/common.cuh:68-183, src/bfs/bfs.cu:56-153) with three kernels and no
// host round trip:
//   k_count  one block per partition: active vertices and their edges
//   k_scan   one block: exclusive prefixes per queue (host / HBM), totals,
//            partitions touched
//   k_write  one block per partition with work: ordered queue entries
//            (vertex id, first edge in the level's work space) and the
//            chunk-start table the expansion kernel indexes by chunk.
// A host partition with no active vertex produces no entry, so its edges are
// never read (the per-partition activity filter).  Queues come out sorted by
// vertex id, which makes the zero-copy reads of a dense host frontier
// sequential.
#include "internal.hpp"
#include "scan.hpp"

namespace thv {

constexpr int CBS = 256;  // compaction block size

__device__ __forceinline__ bool active_bit(const uint32_t *front, uint32_t v) {
  return (front[v >> 5] >> (v & 31)) & 1u;
}

__global__ void __launch_bounds__(CBS)
k_count(const uint32_t *__restrict__ front, const uint32_t *__restrict__ pv, uint32_t P,
        const uint32_t *__restrict__ deg, uint32_t *__restrict__ pcnt,
        unsigned long long *__restrict__ pdeg) {
  __shared__ unsigned long long sh[CBS / 64];
  for (uint32_t p = blockIdx.x; p < P; p += gridDim.x) {
    const uint32_t vs = pv[p], ve = pv[p + 1];
    uint32_t cnt = 0;
    unsigned long long ds = 0;
    for (uint32_t v = vs + threadIdx.x; v < ve; v += CBS) {
      if (active_bit(front, v)) {
        uint32_t d = deg[v];
        cnt += d != 0;
        ds += d;
      }
    }
    unsigned long long c = block_sum<CBS, unsigned long long>(cnt, sh);
    __syncthreads();
    unsigned long long e = block_sum<CBS, unsigned long long>(ds, sh);
    if (threadIdx.x == 0) {
      pcnt[p] = (uint32_t)c;
      pdeg[p] = e;
    }
    __syncthreads();
  }
}

constexpr int SBS = 1024;

__global__ void __launch_bounds__(SBS)
k_scan(const uint32_t *__restrict__ pcnt, const unsigned long long *__restrict__ pdeg,
       const uint8_t *__restrict__ pres, uint32_t P, uint32_t *__restrict__ pbc,
       unsigned long long *__restrict__ pbe, LevelCtr *ctr) {
  __shared__ unsigned long long sh[4][SBS / 64 + 1];
  const uint32_t per = (P + SBS - 1) / SBS;
  const uint32_t p0 = threadIdx.x * per, p1 = min(P, p0 + per);
  unsigned long long c[2] = {0, 0}, e[2] = {0, 0};
  unsigned int t[2] = {0, 0};
  for (uint32_t p = p0; p < p1; p++) {
    int q = pres[p] ? Q_DEV : Q_HOST;
    c[q] += pcnt[p];
    e[q] += pdeg[p];
    t[q] += pcnt[p] != 0;
  }
  unsigned long long tot[4];
  unsigned long long bc0 = block_excl_scan<SBS>(c[0], sh[0], &tot[0]);
  unsigned long long bc1 = block_excl_scan<SBS>(c[1], sh[1], &tot[1]);
  unsigned long long be0 = block_excl_scan<SBS>(e[0], sh[2], &tot[2]);
  unsigned long long be1 = block_excl_scan<SBS>(e[1], sh[3], &tot[3]);
  __syncthreads();
  unsigned long long tt = block_sum<SBS, unsigned long long>(
      (unsigned long long)t[0] | ((unsigned long long)t[1] << 32), sh[0]);
  unsigned long long bc[2] = {bc0, bc1}, be[2] = {be0, be1};
  for (uint32_t p = p0; p < p1; p++) {
    int q = pres[p] ? Q_DEV : Q_HOST;
    pbc[p] = (uint32_t)bc[q];
    pbe[p] = be[q];
    bc[q] += pcnt[p];
    be[q] += pdeg[p];
  }
  if (threadIdx.x == 0) {
    ctr->qn[0] = tot[0];
    ctr->qn[1] = tot[1];
    ctr->qw[0] = tot[2];
    ctr->qw[1] = tot[3];
    ctr->touched[0] = (unsigned)(tt & 0xFFFFFFFFu);
    ctr->touched[1] = (unsigned)(tt >> 32);
  }
}

__global__ void __launch_bounds__(CBS)
k_write(const uint32_t *__restrict__ front, const uint32_t *__restrict__ pv, uint32_t P,
        const uint32_t *__restrict__ deg, const uint8_t *__restrict__ pres,
        const uint32_t *__restrict__ pcnt, const uint32_t *__restrict__ pbc,
        const unsigned long long *__restrict__ pbe, uint32_t *qv0, unsigned long long *qe0,
        uint32_t *qc0, uint32_t *qv1, unsigned long long *qe1, uint32_t *qc1, uint32_t C) {
  __shared__ unsigned long long shc[CBS / 64 + 1], shd[CBS / 64 + 1];
  for (uint32_t p = blockIdx.x; p < P; p += gridDim.x) {
    if (pcnt[p] == 0) continue;
    const uint32_t vs = pv[p], ve = pv[p + 1];
    const bool dev = pres[p] != 0;
    uint32_t *qv = dev ? qv1 : qv0;
    unsigned long long *qe = dev ? qe1 : qe0;
    uint32_t *qc = dev ? qc1 : qc0;
    unsigned long long bc = pbc[p], be = pbe[p];
    for (uint32_t r0 = vs; r0 < ve; r0 += CBS) {
      const uint32_t v = r0 + threadIdx.x;
      uint32_t d = 0;
      if (v < ve && active_bit(front, v)) d = deg[v];
      const bool act = d != 0;
      if (!__syncthreads_or(act)) continue;
      unsigned long long tc, td;
      unsigned long long ic = block_excl_scan<CBS, unsigned long long>(act ? 1ull : 0ull, shc, &tc);
      unsigned long long id = block_excl_scan<CBS, unsigned long long>(d, shd, &td);
      if (act) {
        const unsigned long long idx = bc + ic, e = be + id;
        qv[idx] = v;
        qe[idx] = e;
        for (unsigned long long c = (e + C - 1) / C; c * C < e + d; c++) qc[c] = (uint32_t)idx;
      }
      bc += tc;
      be += td;
      __syncthreads();
    }
  }
}

void launch_compact(Graph &g, QueueSet &w, const uint32_t *front, LevelCtr *ctr,
                    hipStream_t s) {
  Store &st = g.store;
  const uint32_t P = st.P;
  if (P == 0) return;
  const uint32_t grid = P < 65536u ? P : 65536u;
  hipLaunchKernelGGL(k_count, dim3(grid), dim3(CBS), 0, s, front, st.d_pv.as<uint32_t>(), P,
                     g.d_deg.as<uint32_t>(), w.pcnt.as<uint32_t>(),
                     w.pdeg.as<unsigned long long>());
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(SBS), 0, s, w.pcnt.as<uint32_t>(),
                     w.pdeg.as<unsigned long long>(), st.d_pres.as<uint8_t>(), P,
                     w.pbc.as<uint32_t>(), w.pbe.as<unsigned long long>(), ctr);
  hipLaunchKernelGGL(k_write, dim3(grid), dim3(CBS), 0, s, front, st.d_pv.as<uint32_t>(), P,
                     g.d_deg.as<uint32_t>(), st.d_pres.as<uint8_t>(), w.pcnt.as<uint32_t>(),
                     w.pbc.as<uint32_t>(), w.pbe.as<unsigned long long>(),
                     w.qv[0].as<uint32_t>(), w.qe[0].as<unsigned long long>(),
                     w.qc[0].as<uint32_t>(), w.qv[1].as<uint32_t>(),
                     w.qe[1].as<unsigned long long>(), w.qc[1].as<uint32_t>(), w.C);
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
