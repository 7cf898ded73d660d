// batch.hip -- multi-source BFS / SSSP engine (multi.hpp) and thv_batch.
//
// No reference counterpart: the reference runs one source per process
// (src/main.cu:152-220).  The per-level pipeline reuses the split store and
// the partition compaction of the single-source path; only the expansion
// kernels and the per-vertex state differ.
#include <algorithm>

#include "chunk.hpp"
#include "multi.hpp"

namespace thv {

namespace {

constexpr int MBS = 256;
constexpr int ACC_HASH = 0, ACC_CNT = 64, ACC_LM = 128, RES_HASH = 129, RES_CNT = 193,
              RES_DEPTH = 257, ACC_LONG = 321, ACC_MU = 322, ACC_SCAN_HOST = 323,
              ACC_SCAN_HBM = 324, ACC_WORDS = 384;
// bottom-up levels: lists up to this length are scanned by one lane, longer
// ones by a group of BU_G lanes reading BU_R records each per step (measured
// on the 64-lane R-MAT 25 batch: 8 / 8 / 2 beat 32 / 16 / 2 by 5%; larger
// steps lose more to reads past the early exit than they save in steps)
constexpr uint32_t BU_SHORT = 8, BU_G = 8, BU_R = 2;
// per-wave LDS buffer of long-list vertices, flushed to the global queue with
// one atomic per flush (an atomic per wave step on one counter serialises)
constexpr uint32_t BU_LQ = 512;

struct MsArgs {
  const uint32_t *qv[2];
  const unsigned long long *qe[2];
  const uint32_t *qc[2];
  const LevelCtr *ctr;
  const uint64_t *eptr;
  const unsigned long long *F;
  unsigned long long *N, *seen;
  uint32_t *nb;
  uint32_t *dist;
  uint32_t *full;
  unsigned long long *acc;
  uint64_t n;
  uint32_t L;
  uint32_t level;
  uint32_t voff[64];
  unsigned long long fmask;  // BFS: lanes this expansion pushes (others run bottom-up)
};

// BFS: lane per edge, lane masks staged per queue entry.
template <int REC, int EPT>
__global__ void __launch_bounds__(MBS) k_ms_bfs(MsArgs a) {
  constexpr uint32_t C = MBS * EPT, NW = C / 32;
  __shared__ uint32_t s_mask[NW], s_pre[NW];
  __shared__ unsigned long long s_base[C + 1];
  __shared__ unsigned long long s_m[C + 1];
  __shared__ unsigned long long s_hash[64];
  __shared__ unsigned int s_cnt[64];
  __shared__ uint32_t s_val[64];
  __shared__ unsigned long long s_lm;
  const uint32_t tid = threadIdx.x;
  if (tid < 64) {
    s_hash[tid] = 0;
    s_cnt[tid] = 0;
    s_val[tid] = a.level + 2 - a.voff[tid];
  }
  if (tid == 0) s_lm = 0;
  const ChunkSrc cs = chunk_src(a.ctr, C);
  for (uint32_t c = blockIdx.x; c < cs.total; c += gridDim.x) {
    int q;
    uint32_t cc;
    chunk_pick(cs, c, q, cc);
    const ChunkInfo ci = chunk_info<C>(cs, a.qc, q, cc);
    for (uint32_t t = tid; t < NW; t += MBS) s_mask[t] = 0;
    __syncthreads();
    for (uint32_t t = tid; t < ci.nent; t += MBS) {
      const uint32_t v = a.qv[q][ci.i0 + t];
      const unsigned long long b = a.qe[q][ci.i0 + t];
      const unsigned long long s = chunk_rel(b, ci.e0);
      if (s < C) {
        atomicOr(&s_mask[s >> 5], 1u << (s & 31));
        s_base[t] = a.eptr[v] - b * REC;
        s_m[t] = a.F[v] & a.fmask;
      }
    }
    __syncthreads();
    chunk_prefix<NW>(s_mask, s_pre);
    __syncthreads();
    uint32_t dst[EPT], ent[EPT];
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      const uint32_t l = tid + j * MBS;
      dst[j] = 0xFFFFFFFFu;
      ent[j] = 0;
      if (l < ci.len) {
        const uint32_t k = chunk_rank(s_mask, s_pre, l);
        ent[j] = k;
        dst[j] = *reinterpret_cast<const uint32_t *>(s_base[k] + (ci.e0 + l) * REC);
      }
    }
    // phases over all EPT edges, each issuing its memory operations before
    // any result is used: the seen masks (an absent edge reads entry 0 and
    // is ignored), the seen atomics, the next-frontier atomics
    unsigned long long sw[EPT], add[EPT], gain[EPT], on[EPT];
#pragma unroll
    for (int j = 0; j < EPT; j++) sw[j] = a.seen[dst[j] != 0xFFFFFFFFu ? dst[j] : 0];
#pragma unroll
    for (int j = 0; j < EPT; j++) add[j] = dst[j] != 0xFFFFFFFFu ? s_m[ent[j]] & ~sw[j] : 0ull;
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      sw[j] = ~0ull;
      if (add[j]) sw[j] = atomicOr(&a.seen[dst[j]], add[j]);
    }
#pragma unroll
    for (int j = 0; j < EPT; j++) gain[j] = add[j] & ~sw[j];
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      on[j] = ~0ull;
      if (gain[j]) on[j] = atomicOr(&a.N[dst[j]], gain[j]);
    }
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      const uint32_t w = dst[j];
      const unsigned long long gained = gain[j];
      if (!gained) continue;
      if (on[j] == 0ull) atomicOr(&a.nb[w >> 5], 1u << (w & 31));
      atomicOr(&s_lm, gained);
      unsigned long long g = gained;
      while (g) {
        const int qq = __ffsll((long long)g) - 1;
        g &= g - 1;
        const uint32_t val = s_val[qq];
        atomicAdd(&s_hash[qq], mix64(((unsigned long long)w << 32) | val));
        atomicAdd(&s_cnt[qq], 1u);
        if (a.full) a.full[(uint64_t)qq * a.n + w] = val;
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (tid < 64 && s_cnt[tid]) {
    atomicAdd(&a.acc[ACC_HASH + tid], s_hash[tid]);
    atomicAdd(&a.acc[ACC_CNT + tid], (unsigned long long)s_cnt[tid]);
  }
  if (tid == 0 && s_lm) atomicOr(&a.acc[ACC_LM], s_lm);
}

// ---- bottom-up BFS levels (symmetric graphs) ---------------------------------
//
// Every vertex w that some active lane has not reached looks for those lanes
// among its neighbours' frontier masks: N[w] = (OR of F[v] over the
// neighbours v) & ~seen[w], and the scan stops as soon as all of them are
// found.  On a symmetric graph w's out-list is its in-list, so the lanes
// found are exactly those a top-down level hands w.  Each w has one owner
// (a lane, or a group for a long list), so N, seen and the levels need no
// atomics; in the dense middle levels most vertices find all their missing
// lanes within a few records.
struct MsBuArgs {
  const uint64_t *eptr;
  const uint32_t *deg;
  const uint32_t *rbit;
  const uint32_t *fb;           // frontier bitmap: F[v] is read only when v's bit is set
  const unsigned long long *F;
  unsigned long long *N, *seen;
  uint32_t *nb;
  uint32_t *full;
  unsigned long long *acc;
  uint32_t *lq;                 // vertices with long lists, for k_ms_bu_long
  uint64_t n;
  unsigned long long act;       // active lanes
  uint32_t level;
  uint32_t uval;                // the level value when every active lane shares it, else 0
  uint32_t voff[64];
};

// Per-block digest state for bottom-up levels: lane q of every wave keeps
// query q's sum and count in registers.
struct BuDigest {
  unsigned long long h = 0, scan_host = 0, scan_hbm = 0, mu = 0, lm = 0;
  uint32_t cnt = 0;
};

__device__ __forceinline__ unsigned long long readlane64(unsigned long long x, int i) {
  return ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), i)
          << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, i);
}

// The wave's gains, one vertex per set bit of `who` (a lane mask): lane q
// adds query q's pair (w, value) for each vertex whose gain holds q.  When
// all active lanes share one value (uval != 0), the vertex's owner lane has
// hashed the pair already (hw) and the other lanes only add it.
__device__ __forceinline__ void bu_digest_wave(unsigned long long who, unsigned long long gain,
                                               uint32_t w, unsigned long long hw, uint32_t uval,
                                               uint32_t myval, uint32_t lane, uint32_t *full,
                                               uint64_t n, BuDigest &d) {
  while (who) {
    const int i = __builtin_ctzll(who);
    who &= who - 1;
    const unsigned long long gi = readlane64(gain, i);
    d.lm |= gi;
    if (uval) {
      const unsigned long long hi = readlane64(hw, i);
      if ((gi >> lane) & 1ull) {
        d.h += hi;
        d.cnt++;
        if (full) full[(uint64_t)lane * n + (uint32_t)__builtin_amdgcn_readlane((int)w, i)] = uval;
      }
    } else {
      const uint32_t wi = (uint32_t)__builtin_amdgcn_readlane((int)w, i);
      if ((gi >> lane) & 1ull) {
        d.h += mix64(((unsigned long long)wi << 32) | myval);
        d.cnt++;
        if (full) full[(uint64_t)lane * n + wi] = myval;
      }
    }
  }
}

__device__ __forceinline__ void bu_digest_flush(BuDigest &d, unsigned long long *acc,
                                                unsigned long long *s_hash, unsigned int *s_cnt,
                                                unsigned long long *s_misc) {
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  if (d.cnt) {
    atomicAdd(&s_hash[lane], d.h);
    atomicAdd(&s_cnt[lane], d.cnt);
  }
  if (d.lm) atomicOr(&s_misc[0], d.lm);
  if (d.mu) atomicAdd(&s_misc[1], d.mu);
  if (d.scan_host) atomicAdd(&s_misc[2], d.scan_host);
  if (d.scan_hbm) atomicAdd(&s_misc[3], d.scan_hbm);
  __syncthreads();
  if (tid < 64 && s_cnt[tid]) {
    atomicAdd(&acc[ACC_HASH + tid], s_hash[tid]);
    atomicAdd(&acc[ACC_CNT + tid], (unsigned long long)s_cnt[tid]);
  }
  if (tid == 0) {
    if (s_misc[0]) atomicOr(&acc[ACC_LM], s_misc[0]);
    if (s_misc[1]) atomicAdd(&acc[ACC_MU], s_misc[1]);
    if (s_misc[2]) atomicAdd(&acc[ACC_SCAN_HOST], s_misc[2]);
    if (s_misc[3]) atomicAdd(&acc[ACC_SCAN_HBM], s_misc[3]);
  }
}

// Copy a wave's LDS buffer of `nl` long-list vertices to the global queue.
__device__ __forceinline__ void bu_lq_flush(const uint32_t *buf, uint32_t nl,
                                            unsigned long long *acc, uint32_t *lq,
                                            uint32_t lane) {
  // the buffer was written by other lanes of this wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint32_t base = 0;
  if (lane == 0) base = (uint32_t)atomicAdd(&acc[ACC_LONG], (unsigned long long)nl);
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
  for (uint32_t i = lane; i < nl; i += 64) lq[base + i] = buf[i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sweep: one wave per 64 consecutive vertices.  Lists of up to BU_SHORT
// records are scanned by their own lane (4 records and 4 frontier masks in
// flight per step); longer ones are queued for k_ms_bu_long.  The next-
// frontier bitmap words come from a ballot, one plain store per half wave.
template <int REC>
__global__ void __launch_bounds__(MBS) k_ms_bu(MsBuArgs a) {
  __shared__ unsigned long long s_hash[64];
  __shared__ unsigned int s_cnt[64];
  __shared__ unsigned long long s_misc[4];
  __shared__ uint32_t s_lq[MBS / 64][BU_LQ];
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  uint32_t *my_lq = s_lq[tid >> 6];
  uint32_t nl = 0;  // entries in my_lq (wave-uniform)
  if (tid < 64) {
    s_hash[tid] = 0;
    s_cnt[tid] = 0;
  }
  if (tid < 4) s_misc[tid] = 0;
  __syncthreads();
  const uint32_t myval = a.level + 2 - a.voff[lane];
  BuDigest dg;
  const uint64_t nwaves = (uint64_t)gridDim.x * (MBS / 64);
  for (uint64_t w0 = ((uint64_t)blockIdx.x * (MBS / 64) + (tid >> 6)) * 64; w0 < a.n;
       w0 += nwaves * 64) {
    const uint64_t w = w0 + lane;
    uint32_t d = 0;
    unsigned long long sw = 0, U = 0;
    if (w < a.n) {
      d = a.deg[w];
      sw = a.seen[w];
    }
    if (d) U = a.act & ~sw;
    const bool lng = U && d > BU_SHORT;
    unsigned long long acc = 0;
    uint32_t scanned = 0;
    if (U && !lng) {
      const unsigned long long b = a.eptr[w];
      for (uint32_t k = 0; k < d; k += 4) {
        uint32_t v[4];
        unsigned long long f[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          v[j] = 0xFFFFFFFFu;
          if (k + j < d) v[j] = *reinterpret_cast<const uint32_t *>(b + (uint64_t)(k + j) * REC);
        }
        uint32_t fw[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          fw[j] = 0;
          if (v[j] != 0xFFFFFFFFu) fw[j] = a.fb[v[j] >> 5];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
          f[j] = 0;
          if ((fw[j] >> (v[j] & 31)) & 1u) f[j] = a.F[v[j]];
        }
        acc |= f[0] | f[1] | f[2] | f[3];
        scanned = min(d, k + 4);
        if ((acc & U) == U) break;
      }
    }
    // queue the long lists through the wave's LDS buffer
    const unsigned long long lb = __ballot(lng);
    if (lb) {
      if (lng) my_lq[nl + __popcll(lb & ((1ull << lane) - 1))] = (uint32_t)w;
      nl += (uint32_t)__popcll(lb);
      if (nl > BU_LQ - 64) {
        bu_lq_flush(my_lq, nl, a.acc, a.lq, lane);
        nl = 0;
      }
    }
    const unsigned long long gain = acc & U;
    if (gain) {
      a.N[w] = gain;
      a.seen[w] = sw | gain;
    }
    if (!lng && (U & ~gain)) dg.mu += d;
    if (scanned) {
      const bool hbm = (a.rbit[w >> 5] >> (w & 31)) & 1u;
      if (hbm) dg.scan_hbm += scanned;
      else dg.scan_host += scanned;
    }
    const unsigned long long gb = __ballot(gain != 0);
    if (lane == 0 && (uint32_t)gb) a.nb[w0 >> 5] = (uint32_t)gb;
    if (lane == 32 && (gb >> 32)) a.nb[(w0 >> 5) + 1] = (uint32_t)(gb >> 32);
    unsigned long long hw = 0;
    if (a.uval && gain) hw = mix64((w << 32) | a.uval);
    bu_digest_wave(gb, gain, (uint32_t)w, hw, a.uval, myval, lane, a.full, a.n, dg);
  }
  if (nl) bu_lq_flush(my_lq, nl, a.acc, a.lq, lane);
  bu_digest_flush(dg, a.acc, s_hash, s_cnt, s_misc);
}

// Long lists: a group of BU_G lanes per queued vertex reads BU_G * BU_R
// records per step, ORs their frontier masks across the group and stops
// when every missing lane is found.
template <int REC>
__global__ void __launch_bounds__(MBS) k_ms_bu_long(MsBuArgs a) {
  __shared__ unsigned long long s_hash[64];
  __shared__ unsigned int s_cnt[64];
  __shared__ unsigned long long s_misc[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63, sub = tid & (BU_G - 1);
  if (tid < 64) {
    s_hash[tid] = 0;
    s_cnt[tid] = 0;
  }
  if (tid < 4) s_misc[tid] = 0;
  __syncthreads();
  const uint32_t myval = a.level + 2 - a.voff[lane];
  const uint64_t total = a.acc[ACC_LONG];
  constexpr uint32_t GPW = 64 / BU_G;  // groups per wave
  BuDigest dg;
  const uint64_t wave = (uint64_t)blockIdx.x * (MBS / 64) + (tid >> 6);
  const uint64_t nwaves = (uint64_t)gridDim.x * (MBS / 64);
  for (uint64_t i0 = wave * GPW; i0 < total; i0 += nwaves * GPW) {
    const uint64_t i = i0 + (lane / BU_G);
    uint32_t w = 0;
    unsigned long long U = 0, sw = 0, acc = 0;
    uint32_t d = 0, scanned = 0;
    if (i < total) {
      w = a.lq[i];
      d = a.deg[w];
      sw = a.seen[w];
      U = a.act & ~sw;
      const unsigned long long b = a.eptr[w];
      for (uint32_t k = 0; k < d; k += BU_G * BU_R) {
        uint32_t v[BU_R];
        unsigned long long f[BU_R];
#pragma unroll
        for (uint32_t r = 0; r < BU_R; r++) {
          const uint32_t e = k + r * BU_G + sub;
          v[r] = 0xFFFFFFFFu;
          if (e < d) v[r] = *reinterpret_cast<const uint32_t *>(b + (uint64_t)e * REC);
        }
        uint32_t fw[BU_R];
#pragma unroll
        for (uint32_t r = 0; r < BU_R; r++) {
          fw[r] = 0;
          if (v[r] != 0xFFFFFFFFu) fw[r] = a.fb[v[r] >> 5];
        }
#pragma unroll
        for (uint32_t r = 0; r < BU_R; r++) {
          f[r] = 0;
          if ((fw[r] >> (v[r] & 31)) & 1u) f[r] = a.F[v[r]];
        }
        unsigned long long x = f[0];
#pragma unroll
        for (uint32_t r = 1; r < BU_R; r++) x |= f[r];
#pragma unroll
        for (uint32_t o = BU_G / 2; o; o >>= 1) x |= __shfl_xor(x, o, BU_G);
        acc |= x;
        scanned = min(d, k + BU_G * BU_R);
        if ((acc & U) == U) break;
      }
    }
    const bool leader = sub == 0 && i < total;
    const unsigned long long gain = acc & U;
    if (leader) {
      if (gain) {
        a.N[w] = gain;
        a.seen[w] = sw | gain;
        atomicOr(&a.nb[w >> 5], 1u << (w & 31));
      }
      if (U & ~gain) dg.mu += d;
      const bool hbm = (a.rbit[w >> 5] >> (w & 31)) & 1u;
      if (hbm) dg.scan_hbm += scanned;
      else dg.scan_host += scanned;
    }
    unsigned long long hw = 0;
    if (a.uval && leader && gain) hw = mix64(((unsigned long long)w << 32) | a.uval);
    bu_digest_wave(__ballot(leader && gain != 0), gain, w, hw, a.uval, myval, lane, a.full, a.n,
                   dg);
  }
  bu_digest_flush(dg, a.acc, s_hash, s_cnt, s_misc);
}

// SSSP: lane q of a wave relaxes query q.  A wave owns 64 consecutive
// work-space edges at a time: each lane loads one record (coalesced) and
// ranks it once; the wave then walks the 64 edges, broadcasting each with
// readlane, with the source and destination distance rows of 8 edges in
// flight before the first compare.
constexpr uint32_t SSSP_C = 1024;  // edges per chunk (4 waves x 4 x 64)

__global__ void __launch_bounds__(MBS) k_ms_sssp(MsArgs a) {
  constexpr uint32_t C = SSSP_C, NW = C / 32;
  constexpr uint32_t WAVES = MBS / 64, PER_WAVE = C / WAVES;
  __shared__ uint32_t s_mask[NW], s_pre[NW];
  __shared__ unsigned long long s_base[C + 1];
  __shared__ unsigned long long s_m[C + 1];
  __shared__ uint32_t s_v[C + 1];
  __shared__ unsigned long long s_lm;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint32_t L = a.L;
  const bool lane_ok = lane < L;
  if (tid == 0) s_lm = 0;
  const ChunkSrc cs = chunk_src(a.ctr, C);
  for (uint32_t c = blockIdx.x; c < cs.total; c += gridDim.x) {
    int q;
    uint32_t cc;
    chunk_pick(cs, c, q, cc);
    const ChunkInfo ci = chunk_info<C>(cs, a.qc, q, cc);
    for (uint32_t t = tid; t < NW; t += MBS) s_mask[t] = 0;
    __syncthreads();
    for (uint32_t t = tid; t < ci.nent; t += MBS) {
      const uint32_t v = a.qv[q][ci.i0 + t];
      const unsigned long long b = a.qe[q][ci.i0 + t];
      const unsigned long long s = chunk_rel(b, ci.e0);
      if (s < C) {
        atomicOr(&s_mask[s >> 5], 1u << (s & 31));
        s_base[t] = a.eptr[v] - b * 8;
        s_m[t] = a.F[v];
        s_v[t] = v;
      }
    }
    __syncthreads();
    chunk_prefix<NW>(s_mask, s_pre);
    __syncthreads();
    for (uint32_t g0 = wv * PER_WAVE; g0 < (wv + 1) * PER_WAVE && g0 < ci.len; g0 += 64) {
      const uint32_t cnt = min(64u, ci.len - g0);
      // lane j holds edge g0 + j: destination, weight, entry
      uint32_t my_w = 0, my_wt = 0, my_k = 0;
      if (lane < cnt) {
        my_k = chunk_rank(s_mask, s_pre, g0 + lane);
        const uint2 r = *reinterpret_cast<const uint2 *>(s_base[my_k] + (ci.e0 + g0 + lane) * 8);
        my_w = r.x;
        my_wt = r.y;
      }
      // consecutive edges mostly share their source: its lane mask and
      // distance row are read once per source, not once per edge (a stale
      // row is still a valid distance; an improved source is in N anyway)
      uint32_t kprev = 0xFFFFFFFFu, dprev = 0;
      bool mprev = false;
      for (uint32_t j0 = 0; j0 < cnt; j0 += 8) {
        uint32_t w[8], dv[8], dw[8], wt[8];
        bool on[8];
        uint32_t kk[8], ds[8];
        bool nw[8], ms[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const uint32_t jj = min(j0 + j, cnt - 1);
          w[j] = __builtin_amdgcn_readlane(my_w, jj);
          wt[j] = __builtin_amdgcn_readlane(my_wt, jj);
          kk[j] = __builtin_amdgcn_readlane(my_k, jj);
          nw[j] = kk[j] != (j ? kk[j - 1] : kprev);  // a new source (uniform)
        }
        // the new sources' lane masks, then their rows, all in flight together
#pragma unroll
        for (int j = 0; j < 8; j++) {
          ms[j] = false;
          if (nw[j]) ms[j] = lane_ok && ((s_m[kk[j]] >> lane) & 1ull);
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
          ds[j] = 0u;
          if (nw[j] && ms[j]) ds[j] = a.dist[(uint64_t)s_v[kk[j]] * L + lane];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
          if (nw[j]) {
            mprev = ms[j];
            dprev = ds[j];
          }
          on[j] = j0 + j < cnt && mprev;
          dv[j] = dprev;
        }
        kprev = kk[7];
#pragma unroll
        for (int j = 0; j < 8; j++) dw[j] = on[j] ? a.dist[(uint64_t)w[j] * L + lane] : 0u;
        // An improvement is judged against the row just read, so the
        // atomicMin needs no result and the 8 edges' atomics stay in flight
        // together.  A lane whose atomicMin loses to a smaller concurrent
        // write is marked anyway: that write lowered dist[w] in this level
        // and marks the same (w, lane), so the next frontier is unchanged.
        unsigned long long b[8], old[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          bool imp = false;
          if (on[j]) {
            const unsigned long long nd = (unsigned long long)dv[j] + wt[j];
            if (nd < dw[j]) {
              imp = true;
              atomicMin(&a.dist[(uint64_t)w[j] * L + lane], (uint32_t)nd);
            }
          }
          b[j] = __ballot(imp);
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
          old[j] = ~0ull;
          if (b[j] && lane == 0) old[j] = atomicOr(&a.N[w[j]], b[j]);
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
          if (b[j] && lane == 0) {
            if (old[j] == 0ull) atomicOr(&a.nb[w[j] >> 5], 1u << (w[j] & 31));
            atomicOr(&s_lm, b[j]);
          }
        }
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (tid == 0 && s_lm) atomicOr(&a.acc[ACC_LM], s_lm);
}

struct AdmitArgs {
  uint32_t count;
  uint32_t lane[64];
  uint32_t src[64];
};

__global__ void k_ms_admit(AdmitArgs ad, int algo, unsigned long long *F, uint32_t *fb,
                           unsigned long long *seen, uint32_t *dist, uint32_t L,
                           uint32_t *full, uint64_t n, unsigned long long *acc) {
  const uint32_t i = threadIdx.x;
  if (i >= ad.count) return;
  const uint32_t q = ad.lane[i], s = ad.src[i];
  const unsigned long long bit = 1ull << q;
  if (algo == THV_ALGO_BFS) {
    atomicOr(&seen[s], bit);
    acc[ACC_HASH + q] = mix64(((unsigned long long)s << 32) | 1u);
    acc[ACC_CNT + q] = 1;
    if (full) full[(uint64_t)q * n + s] = 1;
  } else {
    dist[(uint64_t)s * L + q] = 1;
  }
  atomicOr(&F[s], bit);
  atomicOr(&fb[s >> 5], 1u << (s & 31));
}

// Frontier bitmap of some lanes: bit v is set when F[v] holds one of them.
// A wave takes 64 vertices, its ballot is their two bitmap words.
__global__ void __launch_bounds__(MBS) k_ms_lanebits(const unsigned long long *F, uint64_t n,
                                                     unsigned long long lanes, uint32_t *bits) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (MBS / 64);
  for (uint64_t w0 = ((uint64_t)blockIdx.x * (MBS / 64) + (threadIdx.x >> 6)) * 64; w0 < n;
       w0 += nw * 64) {
    const uint64_t v = w0 + lane;
    const unsigned long long b = __ballot(v < n && (F[v] & lanes) != 0);
    if (lane < 2) bits[(w0 >> 5) + lane] = (uint32_t)(b >> (32 * lane));
  }
}

__global__ void k_ms_clear(unsigned long long *seen, uint64_t n, unsigned long long keep) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x)
    seen[v] &= keep;
}

// Digest finished SSSP lanes: lane q of each wave reads column q of a row.
__global__ void __launch_bounds__(MBS)
k_ms_sssp_finish(uint32_t *dist, uint64_t n, uint32_t L, unsigned long long fm, uint32_t inf,
                 bool reset, unsigned long long *acc) {
  __shared__ unsigned long long s_h[MBS / 64][64], s_c[MBS / 64][64];
  __shared__ uint32_t s_d[MBS / 64][64];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool mine = lane < L && ((fm >> lane) & 1ull);
  unsigned long long h = 0, cnt = 0;
  uint32_t dep = 0;
  const uint64_t nw = (uint64_t)gridDim.x * (MBS / 64);
  for (uint64_t v = blockIdx.x * (uint64_t)(MBS / 64) + wv; v < n; v += nw) {
    if (!mine) continue;
    uint32_t *p = &dist[v * L + lane];
    const uint32_t d = *p;
    if (d != inf) {
      h += mix64((v << 32) | d);
      cnt++;
      dep = max(dep, d);
      if (reset) *p = inf;
    }
  }
  s_h[wv][lane] = h;
  s_c[wv][lane] = cnt;
  s_d[wv][lane] = dep;
  __syncthreads();
  if (threadIdx.x < 64 && mine) {
    for (int i = 1; i < MBS / 64; i++) {
      h += s_h[i][lane];
      cnt += s_c[i][lane];
      dep = max(dep, s_d[i][lane]);
    }
    if (cnt) {
      atomicAdd(&acc[RES_HASH + lane], h);
      atomicAdd(&acc[RES_CNT + lane], cnt);
      atomicMax(&acc[RES_DEPTH + lane], (unsigned long long)dep);
    }
  }
}

__global__ void k_ms_gather(const uint32_t *dist, uint64_t n, uint32_t L, uint32_t q,
                            uint32_t *out) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x)
    out[v] = dist[v * L + q];
}

int grid_for(uint64_t n, int cap) {
  return (int)std::max<uint64_t>(1, std::min<uint64_t>((n + MBS - 1) / MBS, (uint64_t)cap));
}

}  // namespace

MultiEngine::MultiEngine(Graph &gr, int al, uint32_t lanes, bool kf)
    : g(gr), algo(al), L(lanes), keep_full(kf) {
  if (algo != THV_ALGO_BFS && algo != THV_ALGO_SSSP)
    fail(THV_EINVAL, "batched queries are BFS or SSSP");
  if (L == 0 || L > MAXL) fail(THV_EINVAL, "lanes must be in [1, 64]");
  if (algo == THV_ALGO_SSSP && g.rec != 8)
    fail(THV_ENOTSUP, "SSSP needs a weighted graph (8-byte records)");
  memset(&stats, 0, sizeof stats);
  mu_est = g.m;
  store_ensure(g);
  HIPC(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  const uint64_t n = g.n, words = bitmap_words(n);
  qs.ensure(g, chunk());
  for (int i = 0; i < 2; i++) {
    fb[i].alloc(words * 4);
    F[i].alloc((n + 1) * 8);
    HIPC(hipMemsetAsync(fb[i].p, 0, words * 4, stream));
    HIPC(hipMemsetAsync(F[i].p, 0, (n + 1) * 8, stream));
  }
  if (algo == THV_ALGO_BFS) {
    seen.alloc((n + 1) * 8);
    HIPC(hipMemsetAsync(seen.p, 0, (n + 1) * 8, stream));
    tbits.alloc(words * 4);
    HIPC(hipMemsetAsync(tbits.p, 0, words * 4, stream));
    if (keep_full) full.alloc((uint64_t)L * n * 4 + 4);
  } else {
    dist.alloc((uint64_t)L * n * 4 + 4);
    if (n) launch_fill_u32(dist.as<uint32_t>(), (uint32_t)(n + 1), (uint64_t)L * n, stream);
    if (keep_full) full.alloc(n * 4 + 4);
  }
  acc.alloc(ACC_WORDS * 8);
  HIPC(hipMemsetAsync(acc.p, 0, ACC_WORDS * 8, stream));
  HIPC(hipHostMalloc((void **)&h_acc, ACC_WORDS * 8, hipHostMallocDefault));
  HIPC(hipHostMalloc((void **)&h_ctr, 2 * sizeof(LevelCtr), hipHostMallocDefault));
  HIPC(hipStreamSynchronize(stream));
}

MultiEngine::~MultiEngine() {
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
  if (h_acc) (void)hipHostFree(h_acc);
  if (h_ctr) (void)hipHostFree(h_ctr);
}

uint64_t MultiEngine::bytes_needed(const Graph &g, int algo, uint32_t lanes, bool keep_full) {
  const uint64_t n = g.n, words = bitmap_words(n);
  uint64_t b = 2 * (words * 4 + (n + 1) * 8);                       // fb, F
  b += 2 * ((n + 1) * 12) + 2 * (g.m / 1024 + 2) * 4;                // queues
  b += (n / TILE + 1) * TILE * 4;                                     // long-list queue
  if (algo == THV_ALGO_BFS)
    b += (n + 1) * 8 + words * 4 + (keep_full ? (uint64_t)lanes * n * 4 : 0);
  else b += (uint64_t)lanes * n * 4 + (keep_full ? n * 4 : 0);
  return b;
}

uint32_t MultiEngine::chunk() const {
  return algo == THV_ALGO_SSSP ? SSSP_C : 256u * (uint32_t)g.shape.ept;
}

uint64_t MultiEngine::free_mask() const {
  const uint64_t all = L == 64 ? ~0ull : ((1ull << L) - 1);
  return all & ~active;
}

int MultiEngine::admit(uint32_t src, uint64_t tag) {
  if (src >= g.n) fail(THV_EINVAL, "source vertex out of range");
  const uint64_t fm = free_mask();
  if (!fm) return -1;
  const int q = __builtin_ctzll(fm);
  active |= 1ull << q;
  src_of[q] = src;
  tag_of[q] = tag;
  admit_q.emplace_back((uint32_t)q, src);
  return q;
}

void MultiEngine::launch_level() {
  const uint64_t n = g.n;
  const uint64_t words = bitmap_words(n);
  launched_expand = false;
  launched_finish = false;
  // finished BFS lanes keep their seen bits until a query is admitted (no
  // kernel reads the bits of an inactive lane); one pass clears them all
  if (pend_clear && !admit_q.empty()) {
    hipLaunchKernelGGL(k_ms_clear, dim3(grid_for(n, 8192)), dim3(MBS), 0, stream,
                       seen.as<unsigned long long>(), n, ~pend_clear);
    pend_clear = 0;
  }
  if (pend_finish) {
    HIPC(hipMemsetAsync(acc.as<unsigned long long>() + RES_HASH, 0, 192 * 8, stream));
    hipLaunchKernelGGL(k_ms_sssp_finish, dim3(grid_for(n, 2048)), dim3(MBS), 0, stream,
                       dist.as<uint32_t>(), n, L, pend_finish, (uint32_t)(n + 1), true,
                       acc.as<unsigned long long>());
    finalizing = pend_finish;
    pend_finish = 0;
    launched_finish = true;
  }
  if (!admit_q.empty()) {
    AdmitArgs ad;
    memset(&ad, 0, sizeof ad);
    for (auto &x : admit_q) {
      const uint32_t q = x.first;
      if (algo == THV_ALGO_BFS && keep_full)
        launch_fill_u32(full.as<uint32_t>() + (uint64_t)q * n, (uint32_t)(n + 1), n, stream);
      ad.lane[ad.count] = q;
      ad.src[ad.count] = x.second;
      ad.count++;
      voff[q] = level;
      last_val[q] = 1;
      fsz[q] = 1;       // the source
      cnt_prev[q] = 1;
    }
    hipLaunchKernelGGL(k_ms_admit, dim3(1), dim3(64), 0, stream, ad, algo,
                       F[cur].as<unsigned long long>(), fb[cur].as<uint32_t>(),
                       seen.as<unsigned long long>(), dist.as<uint32_t>(), L,
                       full.as<uint32_t>(), n, acc.as<unsigned long long>());
    admit_q.clear();
    mu_est = g.m;  // new lanes miss every vertex
  }
  level_bu = false;
  if (active) {
    HIPC(hipMemsetAsync(F[cur ^ 1].p, 0, (n + 1) * 8, stream));
    HIPC(hipMemsetAsync(fb[cur ^ 1].p, 0, words * 4, stream));
    HIPC(hipMemsetAsync(acc.as<unsigned long long>() + ACC_LM, 0, 8, stream));
    LevelCtr *c = qs.ctr.as<LevelCtr>() + (level & 1);
    if (algo == THV_ALGO_BFS && g.sym && g.dir_opt) {
      // direction: the frontier's edges against the edges of the vertices
      // some lane still misses (an upper bound kept since the last
      // bottom-up level; any admission resets it to m)
      launch_compact(g, qs, fb[cur].as<uint32_t>(), c, stream, CMP_COUNT_ONLY);
      HIPC(hipMemcpyAsync(&h_ctr[level & 1], c, sizeof(LevelCtr), hipMemcpyDeviceToHost,
                          stream));
      HIPC(hipStreamSynchronize(stream));
      const uint64_t mf = h_ctr[level & 1].qw[0] + h_ctr[level & 1].qw[1];
      level_bu = BU_ALPHA * mf > mu_est && mf > n / 8;
      // lanes with small frontiers run this level top-down
      uint64_t td = 0;
      if (level_bu) {
        const uint64_t thr = std::max<uint64_t>(1, n / LANE_TD_DIV);
        for (uint32_t q = 0; q < MAXL; q++)
          if (((active >> q) & 1ull) && fsz[q] < thr) td |= 1ull << q;
        if (td == active) level_bu = false;  // every lane's frontier is small
      }
      level_td_lanes = level_bu ? td : 0;
      if (level_bu) {
        launch_bottom_up(active & ~td);
        if (td) {
          hipLaunchKernelGGL(k_ms_lanebits, dim3(grid_for(n, g.shape.batch_bpc * g.cus)), dim3(MBS), 0, stream,
                             F[cur].as<unsigned long long>(), n, td, tbits.as<uint32_t>());
          launch_compact(g, qs, tbits.as<uint32_t>(), c, stream);
          launch_top_down(c, td);
        }
      } else {
        launch_compact(g, qs, fb[cur].as<uint32_t>(), c, stream, CMP_WRITE_ONLY);
        launch_top_down(c);
      }
    } else {
      launch_compact(g, qs, fb[cur].as<uint32_t>(), c, stream);
      launch_top_down(c);
    }
    launched_expand = true;
  }
  HIPC(hipMemcpyAsync(h_acc, acc.p, ACC_WORDS * 8, hipMemcpyDeviceToHost, stream));
}

void MultiEngine::launch_bottom_up(uint64_t lanes) {
  const uint64_t n = g.n;
  HIPC(hipMemsetAsync(acc.as<unsigned long long>() + ACC_LONG, 0, 4 * 8, stream));
  MsBuArgs b;
  memset(&b, 0, sizeof b);
  b.eptr = g.store.d_eptr.as<uint64_t>();
  b.deg = g.d_deg.as<uint32_t>();
  b.rbit = g.store.d_rbit.as<uint32_t>();
  b.fb = fb[cur].as<uint32_t>();
  b.F = F[cur].as<unsigned long long>();
  b.N = F[cur ^ 1].as<unsigned long long>();
  b.seen = seen.as<unsigned long long>();
  b.nb = fb[cur ^ 1].as<uint32_t>();
  b.full = keep_full ? full.as<uint32_t>() : nullptr;
  b.acc = acc.as<unsigned long long>();
  b.lq = qs.slist.as<uint32_t>();
  b.n = n;
  b.act = lanes;
  b.level = level;
  for (uint32_t q = 0; q < MAXL; q++) b.voff[q] = voff[q];
  // one value for all active lanes (a cohort admitted together): hash each
  // gained vertex once
  b.uval = 0;
  for (uint32_t q = 0; q < MAXL; q++) {
    if (!((lanes >> q) & 1ull)) continue;
    const uint32_t v = level + 2 - voff[q];
    if (!b.uval) b.uval = v;
    else if (b.uval != v) {
      b.uval = 0;
      break;
    }
  }
  if (qs.slist.bytes < n * 4) fail(THV_ESTATE, "bottom-up queue too small");
  const int gs = grid_for(n, g.shape.batch_bpc * g.cus), gl = g.shape.batch_bpc * g.cus;
  if (g.rec == 8) {
    hipLaunchKernelGGL(k_ms_bu<8>, dim3(gs), dim3(MBS), 0, stream, b);
    hipLaunchKernelGGL(k_ms_bu_long<8>, dim3(gl), dim3(MBS), 0, stream, b);
  } else {
    hipLaunchKernelGGL(k_ms_bu<4>, dim3(gs), dim3(MBS), 0, stream, b);
    hipLaunchKernelGGL(k_ms_bu_long<4>, dim3(gl), dim3(MBS), 0, stream, b);
  }
  HIPC(hipGetLastError());
}

void MultiEngine::launch_top_down(LevelCtr *c, uint64_t lanes) {
  const uint64_t n = g.n;
  {
    MsArgs a;
    memset(&a, 0, sizeof a);
    for (int i = 0; i < 2; i++) {
      a.qv[i] = qs.qv[i].as<uint32_t>();
      a.qe[i] = qs.qe[i].as<unsigned long long>();
      a.qc[i] = qs.qc[i].as<uint32_t>();
    }
    a.ctr = c;
    a.eptr = g.store.d_eptr.as<uint64_t>();
    a.F = F[cur].as<unsigned long long>();
    a.N = F[cur ^ 1].as<unsigned long long>();
    a.seen = seen.as<unsigned long long>();
    a.nb = fb[cur ^ 1].as<uint32_t>();
    a.dist = dist.as<uint32_t>();
    a.full = algo == THV_ALGO_BFS && keep_full ? full.as<uint32_t>() : nullptr;
    a.acc = acc.as<unsigned long long>();
    a.n = n;
    a.L = L;
    a.level = level;
    a.fmask = lanes;
    for (uint32_t q = 0; q < MAXL; q++) a.voff[q] = voff[q];
    const int grid = g.shape.blocks_per_cu * g.cus;
    const bool r8 = g.rec == 8;
    if (qs.C != chunk()) fail(THV_ESTATE, "queue chunk size mismatch");
    if (algo == THV_ALGO_BFS) {
      if (g.shape.ept == 4) {
        if (r8) hipLaunchKernelGGL((k_ms_bfs<8, 4>), dim3(grid), dim3(MBS), 0, stream, a);
        else hipLaunchKernelGGL((k_ms_bfs<4, 4>), dim3(grid), dim3(MBS), 0, stream, a);
      } else {
        if (r8) hipLaunchKernelGGL((k_ms_bfs<8, 8>), dim3(grid), dim3(MBS), 0, stream, a);
        else hipLaunchKernelGGL((k_ms_bfs<4, 8>), dim3(grid), dim3(MBS), 0, stream, a);
      }
    } else {
      // One block per chunk, not a persistent grid: an SSSP level runs for
      // milliseconds, and short blocks let the work-groups of other streams
      // (a BFS cohort's levels) take CUs as they free up.  The chunk count
      // comes from the compaction just launched (one readback).
      HIPC(hipMemcpyAsync(&h_ctr[level & 1], c, sizeof(LevelCtr), hipMemcpyDeviceToHost,
                          stream));
      HIPC(hipStreamSynchronize(stream));
      const LevelCtr &h = h_ctr[level & 1];
      const uint64_t chunks = (h.qw[0] + SSSP_C - 1) / SSSP_C + (h.qw[1] + SSSP_C - 1) / SSSP_C;
      const uint64_t sg = std::max<uint64_t>(1, std::min<uint64_t>(chunks, 1u << 30));
      hipLaunchKernelGGL(k_ms_sssp, dim3((uint32_t)sg), dim3(MBS), 0, stream, a);
    }
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(&h_ctr[level & 1], c, sizeof(LevelCtr), hipMemcpyDeviceToHost, stream));
  }
}

void MultiEngine::finish_level(std::vector<LaneResult> &done) {
  HIPC(hipStreamSynchronize(stream));
  if (launched_finish) {
    for (uint32_t q = 0; q < L; q++) {
      if (!((finalizing >> q) & 1ull)) continue;
      // the lane may already hold a new query; its finished one was
      // recorded in fin_tag/fin_src when the digest was launched
      LaneResult r;
      r.tag = fin_tag[q];
      r.src = fin_src[q];
      r.reached = h_acc[RES_CNT + q];
      r.hash = h_acc[RES_HASH + q];
      r.depth = (uint32_t)h_acc[RES_DEPTH + q];
      done.push_back(r);
    }
    finalizing = 0;
  }
  just_finished = 0;
  if (!launched_expand) return;
  if (level_bu) {
    const uint64_t sh = h_acc[ACC_SCAN_HOST], sd = h_acc[ACC_SCAN_HBM];
    stats.edges_traversed += sh + sd;
    stats.host_bytes += sh * (uint64_t)g.rec;
    stats.hbm_bytes += sd * (uint64_t)g.rec;
    stats.bottom_up_levels++;
    mu_est = h_acc[ACC_MU];
    if (level_td_lanes) {  // the lanes that ran top-down in this level
      const LevelCtr &h = h_ctr[level & 1];
      stats.edges_traversed += h.qw[0] + h.qw[1];
      stats.host_bytes += h.qw[0] * (uint64_t)g.rec;
      stats.hbm_bytes += h.qw[1] * (uint64_t)g.rec;
      // their missing vertices' edges are not in the bottom-up measure
      mu_est = g.m;
    }
  } else {
    const LevelCtr &h = h_ctr[level & 1];
    stats.edges_traversed += h.qw[0] + h.qw[1];
    stats.host_bytes += h.qw[0] * (uint64_t)g.rec;
    stats.hbm_bytes += h.qw[1] * (uint64_t)g.rec;
  }
  stats.iterations++;
  stats.expand_launches++;
  const uint64_t lm = h_acc[ACC_LM];
  for (uint32_t q = 0; q < MAXL; q++) {  // each lane's next frontier: this level's gains
    if (!((active >> q) & 1ull)) continue;
    fsz[q] = h_acc[ACC_CNT + q] - cnt_prev[q];
    cnt_prev[q] = h_acc[ACC_CNT + q];
  }
  for (uint32_t q = 0; q < L; q++) {
    const uint64_t bit = 1ull << q;
    if (!(active & bit)) continue;
    if (lm & bit) {
      last_val[q] = level + 2 - voff[q];
      continue;
    }
    active &= ~bit;
    just_finished |= bit;
    if (algo == THV_ALGO_BFS) {
      LaneResult r;
      r.tag = tag_of[q];
      r.src = src_of[q];
      r.reached = h_acc[ACC_CNT + q];
      r.hash = h_acc[ACC_HASH + q];
      r.depth = last_val[q];
      done.push_back(r);
      pend_clear |= bit;
    } else {
      pend_finish |= bit;
      fin_tag[q] = tag_of[q];
      fin_src[q] = src_of[q];
    }
  }
  cur ^= 1;
  level++;
}

void MultiEngine::copy_full(uint32_t q, uint32_t *out) {
  if (!keep_full) fail(THV_ESTATE, "engine does not keep value arrays");
  const uint64_t n = g.n;
  if (algo == THV_ALGO_BFS) {
    HIPC(hipMemcpyAsync(out, full.as<uint32_t>() + (uint64_t)q * n, n * 4,
                        hipMemcpyDeviceToHost, stream));
  } else {
    hipLaunchKernelGGL(k_ms_gather, dim3(grid_for(n, 8192)), dim3(MBS), 0, stream,
                       dist.as<uint32_t>(), n, L, q, full.as<uint32_t>());
    HIPC(hipMemcpyAsync(out, full.p, n * 4, hipMemcpyDeviceToHost, stream));
  }
  HIPC(hipStreamSynchronize(stream));
}

void run_batch(Graph &g, int algo, const uint32_t *srcs, uint32_t k, uint64_t *hash,
               uint64_t *reached, uint32_t *depth, uint32_t *full, thv_stats *st) {
  if (k == 0) {
    if (st) memset(st, 0, sizeof *st);
    return;
  }
  for (uint32_t i = 0; i < k; i++)
    if (srcs[i] >= g.n) fail(THV_EINVAL, "source vertex out of range");
  MultiEngine e(g, algo, k, full != nullptr);
  for (uint32_t i = 0; i < k; i++) e.admit(srcs[i], i);
  std::vector<LaneResult> done;
  const uint64_t t0 = now_ns();
  while (e.busy()) {
    e.launch_level();
    e.finish_level(done);
    if (full && e.just_finished) {
      for (uint32_t q = 0; q < k; q++)
        if ((e.just_finished >> q) & 1ull) e.copy_full(q, full + (uint64_t)e.tag_of[q] * g.n);
    }
  }
  if (st) {
    *st = e.stats;
    st->ms = (now_ns() - t0) * 1e-6;
  }
  for (auto &r : done) {
    if (hash) hash[r.tag] = r.hash;
    if (reached) reached[r.tag] = r.reached;
    if (depth) depth[r.tag] = r.depth;
  }
}

}  // namespace thv
