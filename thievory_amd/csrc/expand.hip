// expand.hip -- edge-balanced frontier expansion, the hot kernel.
//
// Replaces the reference's warp-per-vertex edge kernels
// (BFS32_Static/Demand/Filter_Kernel src/bfs/bfs_kernels.cu:4-76, 111-154,
// 196-287; CC32_* src/cc/cc_kernels.cu; SSSP32_* src/sssp/sssp_kernels.cu;
// PR32_* src/pr/pr_kernels.cu) with one kernel per algorithm that walks the
// level's work space -- the concatenated edge lists of the queued vertices --
// in chunks of C = 256 * EPT edges (chunk.hpp), host-resident and HBM
// chunks interleaved.
//
// Per chunk (one 256-thread block):
//   1. the chunk-start table names the first queue entry; entries whose
//      lists intersect the chunk are staged in LDS (at most C + 1), and the
//      chunk-relative start of each is marked in an LDS bit mask;
//   2. one wave turns the mask into per-word prefix popcounts;
//   3. every thread takes EPT edges spaced 256 apart: its entry is the rank
//      of the last mark at or before its position (popcount), its record
//      address is entry_base + edge_index * record_bytes; all EPT loads are
//      issued before any is consumed, then each edge is relaxed.
// Hubs are therefore split across many blocks and low-degree vertices share
// a wave, with consecutive lanes reading consecutive addresses whenever the
// queued vertices are adjacent (coalesced 256-byte wave reads, over the
// host link as well as from HBM).
//
// Relaxation rules (identical fixpoints to the reference):
//   BFS  : first visit wins (atomicOr on the visited bitmap), value = level
//          (bfs_kernels.cu:29-44)
//   CC   : atomicMin(val[w], val[v]) (cc_kernels.cu:50-66)
//   SSSP : atomicMin(val[w], val[v] + weight) (sssp_kernels.cu:124-141)
//   PR   : push atomicAdd(res[w], delta[v]) (pr_kernels.cu:539-542), into a
//          per-block LDS sum first when w is one of the hub destinations;
//          pull sum[v] += contrib[u] over in-edges (pr_kernels.cu:45-51)
//   HINT  (store set-up, not a traversal): key64[v] = max over the edges
//          of (deg[w] << 32 | ~w), reduced per entry in LDS
//   CC link (union-find CC of a symmetric graph, cc.hip): link(v, w)
#include "chunk.hpp"
#include "uf.hpp"

namespace thv {

constexpr int EBS = 256;

// Slot of a hub id in the PR-push LDS table (Fibonacci hashing).
__device__ __forceinline__ uint32_t hub_slot(uint32_t w) {
  static_assert((PR_HUB_SLOTS & (PR_HUB_SLOTS - 1)) == 0, "power-of-two table");
  return (w * 2654435761u) >> (32 - __builtin_ctz(PR_HUB_SLOTS));
}

// BFS relaxation of EPT edges whose destinations are loaded (an absent edge
// is 0xFFFFFFFF), each phase issuing the memory operations of all EPT edges
// before any result is used: the visited words, the returning atomics of
// the unvisited ones, the next-frontier bits; with NEXT, the winners' degree
// and residency words, their levels and the next frontier's sums.
template <int EPT, bool NEXT>
__device__ __forceinline__ void bfs_relax(const ExpandArgs &a, const uint32_t (&dst)[EPT],
                                          unsigned long long (&nfn)[2],
                                          unsigned long long (&nfe)[2]) {
  uint32_t vw[EPT], old[EPT];
#pragma unroll
  for (int j = 0; j < EPT; j++) vw[j] = a.vis[dst[j] != 0xFFFFFFFFu ? dst[j] >> 5 : 0];
#pragma unroll
  for (int j = 0; j < EPT; j++) {
    const uint32_t w = dst[j], bit = 1u << (w & 31);
    old[j] = 0xFFFFFFFFu;
    if (w != 0xFFFFFFFFu && !(vw[j] & bit)) old[j] = atomicOr(&a.vis[w >> 5], bit);
  }
  // the winner marks the next frontier; without NEXT the next level's
  // compaction writes its level, with NEXT the winner writes it and adds to
  // the sums (its degree and residency words loaded together)
#pragma unroll
  for (int j = 0; j < EPT; j++) {
    const uint32_t w = dst[j], bit = 1u << (w & 31);
    if (!(old[j] & bit)) atomicOr(&a.next[w >> 5], bit);
  }
  if (NEXT) {
    uint32_t dw[EPT], rw[EPT];
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      const uint32_t w = dst[j];
      const bool won = !(old[j] & (1u << (w & 31)));
      dw[j] = won ? a.deg[w] : 0u;
      rw[j] = won ? a.rbit[w >> 5] : 0u;
    }
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      const uint32_t w = dst[j];
      if (old[j] & (1u << (w & 31))) continue;
      a.val[w] = a.level_value;
      const uint32_t q = (rw[j] >> (w & 31)) & 1u;
      nfn[q] += 1;
      nfe[q] += dw[j];
    }
  }
}

template <int ALGO, int REC, int EPT>
__global__ void __launch_bounds__(EBS) k_expand(ExpandArgs a) {
  constexpr uint32_t C = EBS * EPT;
  constexpr uint32_t NW = C / 32;
  __shared__ uint32_t s_mask[NW];
  __shared__ uint32_t s_pre[NW];
  __shared__ unsigned long long s_base[C + 1];
  // CC/SSSP: source value; PR pull, hint build, CC link: vertex id (PR push
  // has no use for it).  PR push runs 2 blocks per CU, bound by LDS: a table
  // of 1,024 hubs gave 3 blocks per CU but was slower (83 ms against 78 per
  // expansion at R-MAT 26), the hub atomics weigh more than occupancy.
  constexpr uint32_t US = ALGO == A_PR_PUSH ? 1 : C + 1;
  __shared__ uint32_t s_u[US];
  __shared__ float s_f[C + 1];     // PR push: delta; PR pull: partial sum
  // PR push: per-block sums for the hub destinations (an open-addressing
  // table of the hub ids, built at kernel start, flushed at the end)
  constexpr uint32_t HS = ALGO == A_PR_PUSH ? PR_HUB_SLOTS : 1;
  __shared__ uint32_t s_hk[HS];
  __shared__ float s_hv[HS];

  // A_BFS_NEXT: the discovered vertices' entries and edges per residency
  unsigned long long nfn[2] = {0, 0}, nfe[2] = {0, 0};
  const ChunkSrc cs = chunk_src(a.ctr, C);
  const uint32_t tid = threadIdx.x;
  if (ALGO == A_PR_PUSH) {
    for (uint32_t t = tid; t < HS; t += EBS) {
      s_hk[t] = 0xFFFFFFFFu;
      s_hv[t] = 0.0f;
    }
    __syncthreads();
    for (uint32_t t = tid; t < a.nhubs; t += EBS) {
      const uint32_t w = a.hubs[t];
      uint32_t sl = hub_slot(w);
      while (atomicCAS(&s_hk[sl], 0xFFFFFFFFu, w) != 0xFFFFFFFFu) sl = (sl + 1) & (HS - 1);
    }
    __syncthreads();
  }

  for (uint32_t c = blockIdx.x; c < cs.total; c += gridDim.x) {
    int q;
    uint32_t cc;
    chunk_pick(cs, c, q, cc);
    const ChunkInfo ci = chunk_info<C>(cs, a.qc, q, cc);
    const uint32_t *__restrict__ qv = a.qv[q];
    const unsigned long long *__restrict__ qe = a.qe[q];

    for (uint32_t t = tid; t < NW; t += EBS) s_mask[t] = 0;
    __syncthreads();
    for (uint32_t t = tid; t < ci.nent; t += EBS) {
      const uint32_t v = qv[ci.i0 + t];
      const unsigned long long b = qe[ci.i0 + t];
      const unsigned long long s = chunk_rel(b, ci.e0);
      if (s < C) {
        atomicOr(&s_mask[s >> 5], 1u << (s & 31));
        s_base[t] = a.eptr[v] - b * REC;
        if (ALGO == A_CC || ALGO == A_SSSP) s_u[t] = a.val[v];
        if (ALGO == A_PR_PUSH) s_f[t] = a.pr_src[v];
        if (ALGO == A_PR_PULL) {
          s_u[t] = v;
          s_f[t] = 0.0f;
        }
        if (ALGO == A_HINT || ALGO == A_CC_LINK) s_u[t] = v;
      }
    }
    __syncthreads();
    chunk_prefix<NW>(s_mask, s_pre);
    __syncthreads();

    uint32_t dst[EPT], wt[EPT], ent[EPT];
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      const uint32_t l = tid + j * EBS;
      dst[j] = 0xFFFFFFFFu;
      wt[j] = 0;
      ent[j] = 0;
      if (l < ci.len) {
        const uint32_t k = chunk_rank(s_mask, s_pre, l);
        ent[j] = k;
        const unsigned long long addr = s_base[k] + (ci.e0 + l) * REC;
        if (REC == 8) {
          const uint2 r = *reinterpret_cast<const uint2 *>(addr);
          dst[j] = r.x;
          wt[j] = r.y;
        } else {
          dst[j] = *reinterpret_cast<const uint32_t *>(addr);
        }
      }
    }

    if (ALGO == A_HINT) {
      // every thread's records are in registers: s_base becomes the
      // entries' running maxima
      __syncthreads();
      for (uint32_t t = tid; t < ci.nent; t += EBS) s_base[t] = 0;
      __syncthreads();
    }

    // Relaxation in phases, each issuing the memory operations of all EPT
    // edges before any result is used: the destination-state words (an
    // absent edge reads word 0 and is ignored), then the returning atomics
    // of the edges that need one, then the next-frontier bits.
    if (ALGO == A_BFS || ALGO == A_BFS_NEXT) {
      bfs_relax<EPT, ALGO == A_BFS_NEXT>(a, dst, nfn, nfe);
    } else if (ALGO == A_CC || ALGO == A_SSSP) {
      uint32_t cur[EPT], old[EPT], nv[EPT];
#pragma unroll
      for (int j = 0; j < EPT; j++) cur[j] = a.val[dst[j] != 0xFFFFFFFFu ? dst[j] : 0];
#pragma unroll
      for (int j = 0; j < EPT; j++) {
        const uint32_t w = dst[j];
        const unsigned long long nv64 =
            (unsigned long long)s_u[ent[j]] + (ALGO == A_SSSP ? wt[j] : 0u);
        nv[j] = (uint32_t)nv64;
        old[j] = 0;
        if (w != 0xFFFFFFFFu && nv64 < cur[j]) old[j] = atomicMin(&a.val[w], nv[j]);
      }
      uint32_t nw[EPT];
#pragma unroll
      for (int j = 0; j < EPT; j++) {
        nw[j] = 0xFFFFFFFFu;
        if (nv[j] < old[j]) nw[j] = a.next[dst[j] >> 5];
      }
#pragma unroll
      for (int j = 0; j < EPT; j++) {
        const uint32_t w = dst[j], bit = 1u << (w & 31);
        if (!(nw[j] & bit)) atomicOr(&a.next[w >> 5], bit);
      }
    } else if (ALGO == A_PR_PUSH) {
      // a hub destination adds into the block's LDS table (it is there:
      // every hub was inserted), any other one atomically in HBM
      uint32_t hb[EPT];
#pragma unroll
      for (int j = 0; j < EPT; j++) hb[j] = a.hubbit[dst[j] != 0xFFFFFFFFu ? dst[j] >> 5 : 0];
#pragma unroll
      for (int j = 0; j < EPT; j++) {
        const uint32_t w = dst[j];
        if (w == 0xFFFFFFFFu) continue;
        const float d = s_f[ent[j]];
        if ((hb[j] >> (w & 31)) & 1u) {
          uint32_t sl = hub_slot(w);
          while (s_hk[sl] != w) sl = (sl + 1) & (HS - 1);
          atomicAdd(&s_hv[sl], d);
        } else {
          unsafeAtomicAdd(&a.pr_dst[w], d);
        }
      }
    } else if (ALGO == A_PR_PULL) {  // dst[j] is an in-neighbour of the entry's vertex
      float c[EPT];
#pragma unroll
      for (int j = 0; j < EPT; j++) c[j] = a.pr_src[dst[j] != 0xFFFFFFFFu ? dst[j] : 0];
#pragma unroll
      for (int j = 0; j < EPT; j++)
        if (dst[j] != 0xFFFFFFFFu) atomicAdd(&s_f[ent[j]], c[j]);
    } else if (ALGO == A_CC_LINK) {
#pragma unroll
      for (int j = 0; j < EPT; j++)
        if (dst[j] != 0xFFFFFFFFu) uf_link(s_u[ent[j]], dst[j], a.val, a.err, a.uf_steps);
    } else if (ALGO == A_HINT) {
      uint32_t dw[EPT];
#pragma unroll
      for (int j = 0; j < EPT; j++) dw[j] = a.deg[dst[j] != 0xFFFFFFFFu ? dst[j] : 0];
#pragma unroll
      for (int j = 0; j < EPT; j++)
        if (dst[j] != 0xFFFFFFFFu)
          atomicMax(&s_base[ent[j]], ((unsigned long long)dw[j] << 32) | (0xFFFFFFFFu - dst[j]));
    }
    if (ALGO == A_PR_PULL) {
      __syncthreads();
      for (uint32_t t = tid; t < ci.nent; t += EBS) {
        if (chunk_rel(qe[ci.i0 + t], ci.e0) < C) unsafeAtomicAdd(&a.pr_dst[s_u[t]], s_f[t]);
      }
    }
    if (ALGO == A_HINT) {
      __syncthreads();
      for (uint32_t t = tid; t < ci.nent; t += EBS) {
        if (chunk_rel(qe[ci.i0 + t], ci.e0) < C && s_base[t])
          atomicMax(&a.key64[s_u[t]], s_base[t]);
      }
    }
    __syncthreads();
  }
  if (ALGO == A_PR_PUSH) {
    __syncthreads();
    for (uint32_t t = tid; t < HS; t += EBS)
      if (s_hk[t] != 0xFFFFFFFFu && s_hv[t] != 0.0f) unsafeAtomicAdd(&a.pr_dst[s_hk[t]], s_hv[t]);
  }
  if (ALGO == A_BFS_NEXT) {  // one atomic per counter and block
    __shared__ unsigned long long s_n[EBS / 64];
    const unsigned long long x0 = block_sum<EBS>(nfn[0], s_n), x1 = block_sum<EBS>(nfn[1], s_n);
    const unsigned long long y0 = block_sum<EBS>(nfe[0], s_n), y1 = block_sum<EBS>(nfe[1], s_n);
    if (tid == 0) {
      if (x0) atomicAdd(&a.nctr->qn[0], x0);
      if (x1) atomicAdd(&a.nctr->qn[1], x1);
      if (y0) atomicAdd(&a.nctr->qw[0], y0);
      if (y1) atomicAdd(&a.nctr->qw[1], y1);
    }
  }
}

// The expansion of one vertex's list (a BFS level whose frontier is {v},
// with the next-frontier sums): no queue, chunk table or staging -- block
// b takes records [b * 1024, ...) of the list, four per thread, in
// grid-stride steps (loaded clamped to the list, so all four are in flight
// together).  Block 0 also writes ctr's totals for the one-entry frontier
// (its residency from rbit, its partition's from vpart and pres) and v's
// value, as the queue kernels would.
constexpr int EPT_ONE = 4;
template <int REC>
__global__ void __launch_bounds__(EBS) k_expand_one(ExpandArgs a, uint32_t v, uint32_t d,
                                                    const uint32_t *vpart, const uint8_t *pres,
                                                    LevelCtr *ctr, uint32_t v_value) {
  unsigned long long nfn[2] = {0, 0}, nfe[2] = {0, 0};
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int q = (a.rbit[v >> 5] >> (v & 31)) & 1u ? Q_DEV : Q_HOST;
    const int tq = pres[vpart[v]] ? Q_DEV : Q_HOST;
    LevelCtr t = *ctr;  // keeps cand
    t.qn[q] = 1;
    t.qn[q ^ 1] = 0;
    t.qw[q] = d;
    t.qw[q ^ 1] = 0;
    t.touched[tq] = 1;
    t.touched[tq ^ 1] = 0;
    t.rd[0] = t.rd[1] = 0;
    *ctr = t;
    a.val[v] = v_value;
  }
  const unsigned long long eb = a.eptr[v];
  for (uint64_t i0 = (uint64_t)blockIdx.x * EBS * EPT_ONE; i0 < d;
       i0 += (uint64_t)gridDim.x * EBS * EPT_ONE) {
    uint32_t dst[EPT_ONE];
#pragma unroll
    for (int j = 0; j < EPT_ONE; j++) {
      const uint64_t i = i0 + threadIdx.x + (uint64_t)j * EBS;
      const uint32_t x = *reinterpret_cast<const __attribute__((address_space(1))) uint32_t *>(
          eb + (i < d ? i : d - 1) * REC);
      dst[j] = i < d ? x : 0xFFFFFFFFu;
    }
    bfs_relax<EPT_ONE, true>(a, dst, nfn, nfe);
  }
  __shared__ unsigned long long s_n[EBS / 64];
  const unsigned long long x0 = block_sum<EBS>(nfn[0], s_n), x1 = block_sum<EBS>(nfn[1], s_n);
  const unsigned long long y0 = block_sum<EBS>(nfe[0], s_n), y1 = block_sum<EBS>(nfe[1], s_n);
  if (threadIdx.x == 0) {
    if (x0) atomicAdd(&a.nctr->qn[0], x0);
    if (x1) atomicAdd(&a.nctr->qn[1], x1);
    if (y0) atomicAdd(&a.nctr->qw[0], y0);
    if (y1) atomicAdd(&a.nctr->qw[1], y1);
  }
}

void launch_expand_one(Graph &g, const ExpandArgs &args, uint32_t v, LevelCtr *ctr,
                       uint32_t v_value, hipStream_t s) {
  if (v >= g.n) fail(THV_EINVAL, "vertex out of range");
  const uint64_t d = g.h_off[v + 1] - g.h_off[v];
  if (d == 0 || d > 0xFFFFFFFFull) fail(THV_ESTATE, "one-list expansion of 1..2^32-1 records");
  if (!args.nctr || !args.rbit || !args.deg || !args.val || !args.vis || !args.next)
    fail(THV_ESTATE, "one-list expansion needs nctr, rbit, deg, val and the bitmaps");
  const uint32_t *vpart = g.store.d_vpart.as<uint32_t>();
  const uint8_t *pres = g.store.d_pres.as<uint8_t>();
  const uint64_t blocks = div_up(d, (uint64_t)EBS * EPT_ONE);
  const int grid = (int)std::min<uint64_t>(blocks, (uint64_t)g.shape.blocks_per_cu * g.cus);
  if (g.rec == 8)
    hipLaunchKernelGGL(k_expand_one<8>, dim3(grid), dim3(EBS), 0, s, args, v, (uint32_t)d, vpart,
                       pres, ctr, v_value);
  else
    hipLaunchKernelGGL(k_expand_one<4>, dim3(grid), dim3(EBS), 0, s, args, v, (uint32_t)d, vpart,
                       pres, ctr, v_value);
  HIPC(hipGetLastError());
}

template <int ALGO, int REC>
static void launch_t(Graph &g, const ExpandArgs &args, hipStream_t s) {
  const int grid = g.shape.blocks_per_cu * g.cus;
  switch (g.shape.ept) {
    case 4:
      hipLaunchKernelGGL((k_expand<ALGO, REC, 4>), dim3(grid), dim3(EBS), 0, s, args);
      break;
    case 8:
      hipLaunchKernelGGL((k_expand<ALGO, REC, 8>), dim3(grid), dim3(EBS), 0, s, args);
      break;
    default:
      fail(THV_EINVAL, "unsupported edges-per-thread");
  }
}

void launch_expand(Graph &g, Algo a, const ExpandArgs &args, hipStream_t s) {
  if (args.C != 256u * (uint32_t)g.shape.ept)
    fail(THV_ESTATE, "queue chunk size does not match the launch shape");
  const bool r8 = g.rec == 8;
  switch (a) {
    case A_BFS: r8 ? launch_t<A_BFS, 8>(g, args, s) : launch_t<A_BFS, 4>(g, args, s); break;
    case A_BFS_NEXT:
      if (!args.nctr || !args.rbit || !args.deg || !args.val)
        fail(THV_ESTATE, "next-frontier bookkeeping needs nctr, rbit, deg and val");
      r8 ? launch_t<A_BFS_NEXT, 8>(g, args, s) : launch_t<A_BFS_NEXT, 4>(g, args, s);
      break;
    case A_CC: r8 ? launch_t<A_CC, 8>(g, args, s) : launch_t<A_CC, 4>(g, args, s); break;
    case A_SSSP:
      if (!r8) fail(THV_ENOTSUP, "SSSP needs 8-byte {dst, weight} records");
      launch_t<A_SSSP, 8>(g, args, s);
      break;
    case A_PR_PUSH:
      r8 ? launch_t<A_PR_PUSH, 8>(g, args, s) : launch_t<A_PR_PUSH, 4>(g, args, s);
      break;
    case A_PR_PULL:
      r8 ? launch_t<A_PR_PULL, 8>(g, args, s) : launch_t<A_PR_PULL, 4>(g, args, s);
      break;
    case A_HINT:
      r8 ? launch_t<A_HINT, 8>(g, args, s) : launch_t<A_HINT, 4>(g, args, s);
      break;
    case A_CC_LINK:
      r8 ? launch_t<A_CC_LINK, 8>(g, args, s) : launch_t<A_CC_LINK, 4>(g, args, s);
      break;
  }
  HIPC(hipGetLastError());
}

}  // namespace thv
