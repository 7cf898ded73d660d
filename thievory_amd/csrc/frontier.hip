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
//   k_sscan   one block: exclusive bases of the segments of 64 tiles (the
//             count kernels add their tiles' sums into them), totals
//   k_twrite  one block per tile with work: ordered queue entries (vertex
//             id, first edge in the level's work space), the chunk-start
//             table the expansion kernel indexes by chunk, and a stamp on
//             every partition touched
//   k_touched the stamped partitions, counted per residency
// A host partition with no active vertex produces no entry, so its edges are
// never read (the per-partition activity filter).  Queues come out sorted by
// vertex id, which makes the zero-copy reads of a dense host frontier
// sequential.  A wave reads its sub-tile's bitmap words with one load and
// the degrees of all its selected vertices with one batch of loads, so an
// empty sub-tile costs one load and a few lane shuffles.
#include <algorithm>

#include "internal.hpp"
#include "scan.hpp"
#include "side.hpp"

namespace thv {

namespace {

constexpr int CBS = 256;
constexpr uint32_t SEG = 64;  // tiles per segment of the two-level scan
constexpr uint32_t SUB = TILE / (CBS / 64);  // vertices per wave
constexpr uint32_t ROUNDS = SUB / 64;
static_assert(SUB == PROBE_REGION, "probe regions are the waves' sub-tiles");

struct CompactArgs {
  const uint32_t *front, *rbit, *deg, *vpart;
  const uint8_t *pres;
  uint64_t n;
  uint32_t ntiles;
  uint32_t P;
  uint32_t *tcnt;               // per tile and queue: entries
  unsigned long long *tdeg;     //   and edges
  unsigned long long *ssum;     // per segment of SEG tiles: the 4 sums (atomics)
  unsigned long long *sbase;    //   and their exclusive bases
  uint32_t nseg;
  uint32_t *qv[2];
  unsigned long long *qe[2];
  uint32_t *qc[2];
  uint32_t *ptouch;
  uint32_t stamp;
  LevelCtr *ctr;
  uint32_t C;
  uint32_t invert;     // select vertices whose bit is clear
  uint32_t no_chunks;  // skip the chunk-start table
  uint32_t *val;       // BFS: frontier vertices' level, written by k_tcount
  uint32_t val_value;
  uint32_t dmin, dmax; // only vertices with dmin <= degree <= dmax (dmin >= 1)
  const uint32_t *fval;  // if set: only vertices with fval[v] < fthr
  uint32_t fthr;
  uint32_t *clear;       // if set (write pass): clear the selected vertices' bits here
  uint32_t *minout;      // if set (count pass, with fval): atomicMin of the selected fval
  LevelSide side;        // count pass: k_tcount clears, k_sscan publishes
};

// One wave's 1024-vertex sub-tile, read in two dependent memory round
// trips instead of one per 64 vertices: lanes 0-31 load the sub-tile's 32
// frontier and residency words together, every round takes its word by a
// lane shuffle, then the degrees of all selected vertices of all rounds are
// loaded together.  Round r of lane l is vertex base + 64 r + l.
struct SubTile {
  uint32_t fmin;       // smallest fval of a selected vertex (with fval), else ~0
  uint32_t act;        // bit r: vertex of round r is selected and its degree in range
  uint32_t dev;        // bit r: its edges are in HBM
  uint32_t raw;        // bit r: its frontier bit is set (before any inversion)
  uint32_t d[ROUNDS];  // degree of round r's vertex when selected, else 0
};

__device__ __forceinline__ void load_subtile(const CompactArgs &a, uint64_t base, uint32_t lane,
                                             SubTile &t) {
  const uint64_t w0 = base >> 5;
  uint32_t fw = 0, rw = 0;
  if (lane < 32 && ((w0 + lane) << 5) < a.n) {
    fw = a.front[w0 + lane];
    rw = a.rbit[w0 + lane];
  }
  t.act = t.dev = t.raw = 0;
#pragma unroll
  for (uint32_t r = 0; r < ROUNDS; r++) {
    const int src = (int)(2 * r + (lane >> 5));
    const uint32_t f = (uint32_t)__shfl((int)fw, src, 64);
    const uint32_t rb = (uint32_t)__shfl((int)rw, src, 64);
    const uint32_t bit = 1u << (lane & 31);
    const bool in = base + r * 64 + lane < a.n;
    const bool fb = (f & bit) != 0;
    if (in && (fb != (a.invert != 0))) t.act |= 1u << r;
    if (in && fb) t.raw |= 1u << r;
    if (rb & bit) t.dev |= 1u << r;
  }
#pragma unroll
  for (uint32_t r = 0; r < ROUNDS; r++)
    t.d[r] = ((t.act >> r) & 1u) ? a.deg[base + r * 64 + lane] : 0u;
  uint32_t fv[ROUNDS];
  if (a.fval) {
#pragma unroll
    for (uint32_t r = 0; r < ROUNDS; r++) {
      fv[r] = 0xFFFFFFFFu;
      if ((t.act >> r) & 1u) fv[r] = a.fval[base + r * 64 + lane];
    }
#pragma unroll
    for (uint32_t r = 0; r < ROUNDS; r++)
      if (fv[r] >= a.fthr) t.d[r] = 0;  // fails the value filter
  }
#pragma unroll
  for (uint32_t r = 0; r < ROUNDS; r++) {
    if (t.d[r] < a.dmin || t.d[r] > a.dmax) {
      t.act &= ~(1u << r);
      t.d[r] = 0;
    }
  }
  t.dev &= t.act;
  t.fmin = 0xFFFFFFFFu;
  if (a.fval) {
#pragma unroll
    for (uint32_t r = 0; r < ROUNDS; r++)
      if ((t.act >> r) & 1u) t.fmin = min(t.fmin, fv[r]);
  }
}

// Per-lane totals of a sub-tile: entries and edges per queue.
__device__ __forceinline__ void subtile_sums(const SubTile &t, uint32_t c[2],
                                             unsigned long long e[2]) {
  c[Q_DEV] = __popc(t.dev);
  c[Q_HOST] = __popc(t.act) - c[Q_DEV];
  e[0] = e[1] = 0;
#pragma unroll
  for (uint32_t r = 0; r < ROUNDS; r++) e[(t.dev >> r) & 1u] += t.d[r];
}

__global__ void __launch_bounds__(CBS) k_tcount(CompactArgs a) {
  __shared__ unsigned long long s[CBS / 64][4];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t fmin = 0xFFFFFFFFu;
  side_clear(a.side);
  for (uint32_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const uint64_t base = (uint64_t)tile * TILE + (uint64_t)wv * SUB;
    unsigned long long v0 = 0, v1 = 0, v2 = 0, v3 = 0;
    // fast path: no word of the sub-tile selects anything
    const uint64_t w0 = base >> 5;
    const uint32_t fw = lane < 32 && ((w0 + lane) << 5) < a.n ? a.front[w0 + lane]
                                                              : (a.invert ? 0xFFFFFFFFu : 0u);
    if (__ballot(a.invert ? fw != 0xFFFFFFFFu : fw != 0u)) {
      SubTile t;
      load_subtile(a, base, lane, t);
      // BFS: every frontier vertex (degree 0 included) takes the level
      // here, in vertex order, instead of by a scattered store when found
      if (a.val) {
#pragma unroll
        for (uint32_t r = 0; r < ROUNDS; r++)
          if ((t.raw >> r) & 1u) a.val[base + r * 64 + lane] = a.val_value;
      }
      fmin = min(fmin, t.fmin);
      uint32_t c[2];
      unsigned long long e[2];
      subtile_sums(t, c, e);
      v0 = wave_sum(c[0]);
      v1 = wave_sum(c[1]);
      v2 = wave_sum(e[0]);
      v3 = wave_sum(e[1]);
    }
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
      if (threadIdx.x < 2) a.tcnt[2 * tile + threadIdx.x] = (uint32_t)x;
      else a.tdeg[2 * tile + threadIdx.x - 2] = x;
      if (x) atomicAdd(&a.ssum[(tile / SEG) * 4 + threadIdx.x], x);
    }
    __syncthreads();
  }
  if (a.minout) {
    // one atomic per wave that holds a smaller value than the one stored
    // (all waves hitting one address serialise; the stored minimum falls
    // fast, so most waves only read it)
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) fmin = min(fmin, (uint32_t)__shfl_xor((int)fmin, d, 64));
    if (lane == 0 && fmin != 0xFFFFFFFFu &&
        fmin < __hip_atomic_load(a.minout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMin(a.minout, fmin);
  }
}

constexpr int SBS = 1024;
constexpr int SNW = SBS / 64;

// Segment bases: the count kernels add each tile's four sums (entries and
// edges per queue) into its segment of SEG tiles; one block scans the
// segments (nseg <= 16384: 2^32 vertices), writes the totals, resets the
// counters and clears the segment sums for the next count.  A tile's base
// is its segment's base plus the tiles before it in the segment, which the
// write pass adds itself (one coalesced load).
__device__ __forceinline__ void sscan_block(const CompactArgs &a) {
  __shared__ unsigned long long sh[4][SNW + 1];
  unsigned long long carry[4] = {0, 0, 0, 0};
  for (uint32_t b = 0; b < a.nseg; b += SBS) {
    const uint32_t sg = b + threadIdx.x;
    unsigned long long x[4] = {0, 0, 0, 0};
    if (sg < a.nseg) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        x[i] = a.ssum[sg * 4 + i];
        a.ssum[sg * 4 + i] = 0;
      }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      unsigned long long tot;
      const unsigned long long ex = block_excl_scan<SBS>(x[i], sh[i], &tot);
      if (sg < a.nseg) a.sbase[sg * 4 + i] = carry[i] + ex;
      carry[i] += tot;
    }
    __syncthreads();
  }
  __shared__ LevelCtr tot;
  if (threadIdx.x == 0) {
    tot = *a.ctr;  // keeps cand
    tot.qn[0] = carry[0];
    tot.qn[1] = carry[1];
    tot.qw[0] = carry[2];
    tot.qw[1] = carry[3];
    tot.touched[0] = 0;
    tot.touched[1] = 0;
    tot.rd[0] = 0;
    tot.rd[1] = 0;
    *a.ctr = tot;
  }
  if (a.side.pub || a.side.xpub || a.side.zctr) {
    __syncthreads();
    if (threadIdx.x < 64) side_publish(a.side, threadIdx.x, a.ctr, &tot);
  }
}

__global__ void __launch_bounds__(SBS) k_sscan(CompactArgs a) { sscan_block(a); }

// Both tile-table sets of a probe in one launch: block k scans set k.
__global__ void __launch_bounds__(SBS) k_sscan2(CompactArgs a0, CompactArgs a1) {
  sscan_block(blockIdx.x ? a1 : a0);
}

// Base of a tile: its segment's base plus the tiles before it in the
// segment (one load per lane, then a wave sum).  Every lane gets it.
__device__ __forceinline__ void tile_bases(const CompactArgs &a, uint32_t tile, uint32_t lane,
                                           unsigned long long tb[4]) {
  const uint32_t sg = tile / SEG, t0 = sg * SEG, t = t0 + lane;
  const bool before = t < tile;
  unsigned long long x[4];
  x[0] = before ? a.tcnt[2 * t] : 0u;
  x[1] = before ? a.tcnt[2 * t + 1] : 0u;
  x[2] = before ? a.tdeg[2 * t] : 0ull;
  x[3] = before ? a.tdeg[2 * t + 1] : 0ull;
#pragma unroll
  for (int i = 0; i < 4; i++) tb[i] = a.sbase[sg * 4 + i] + wave_sum(x[i]);
}

// The write pass gives each lane a run of VPL consecutive vertices (the
// count pass gives it one vertex per round): the entries a lane writes are
// consecutive, so one wave scan of the lanes' totals places them all, in
// place of a scan per 64 vertices.  It selects exactly the vertices the
// count pass counted.
constexpr uint32_t VPL = SUB / 64;
static_assert(VPL == 16, "a lane's run is half a bitmap word");

__global__ void __launch_bounds__(CBS, 4) k_twrite(CompactArgs a) {
  __shared__ unsigned long long s[CBS / 64][4];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (uint32_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    if (a.tcnt[2 * tile] == 0 && a.tcnt[2 * tile + 1] == 0) continue;
    const uint64_t v0 = (uint64_t)tile * TILE + (uint64_t)wv * SUB + (uint64_t)lane * VPL;
    const uint32_t sh = (uint32_t)(v0 & 31);
    unsigned long long tb[4];
    tile_bases(a, tile, lane, tb);
    uint32_t sel = 0, dev = 0;
    if (v0 < a.n) {
      const uint32_t fw = (a.front[v0 >> 5] >> sh) & 0xFFFFu;
      dev = (a.rbit[v0 >> 5] >> sh) & 0xFFFFu;
      sel = (a.invert ? ~fw : fw) & 0xFFFFu;
      if (a.n - v0 < VPL) sel &= (1u << (a.n - v0)) - 1u;
    }
    uint32_t d[VPL];
#pragma unroll
    for (uint32_t r = 0; r < VPL; r++) d[r] = ((sel >> r) & 1u) ? a.deg[v0 + r] : 0u;
    if (a.fval) {
      uint32_t fv[VPL];
#pragma unroll
      for (uint32_t r = 0; r < VPL; r++) {
        fv[r] = 0;
        if ((sel >> r) & 1u) fv[r] = a.fval[v0 + r];
      }
#pragma unroll
      for (uint32_t r = 0; r < VPL; r++)
        if (fv[r] >= a.fthr) d[r] = 0;  // fails the value filter
    }
#pragma unroll
    for (uint32_t r = 0; r < VPL; r++) {
      if (d[r] < a.dmin || d[r] > a.dmax) {
        sel &= ~(1u << r);
        d[r] = 0;
      }
    }
    dev &= sel;
    // the selected vertices leave the bitmap (two lanes share a word: each
    // clears its own half)
    if (a.clear && sel) atomicAnd(&a.clear[v0 >> 5], ~(sel << sh));
    // partitions of the lane's first and last selected vertices (a run of
    // 16 vertices rarely spans two partitions; then each is looked up)
    uint32_t pf = 0xFFFFFFFFu, pl = 0xFFFFFFFFu;
    if (sel) {
      pf = a.vpart[v0 + (__ffs(sel) - 1)];
      pl = a.vpart[v0 + (31 - __clz(sel))];
    }
    // lane totals -> exclusive offsets within the wave, then the block
    const unsigned long long cd = __popc(dev), ch = __popc(sel) - cd;
    unsigned long long ed = 0, eh = 0;
#pragma unroll
    for (uint32_t r = 0; r < VPL; r++) {
      if ((dev >> r) & 1u) ed += d[r];
      else eh += d[r];
    }
    const unsigned long long ich = wave_incl_scan(ch), icd = wave_incl_scan(cd);
    const unsigned long long ieh = wave_incl_scan(eh), ied = wave_incl_scan(ed);
    if (lane == 63) {
      s[wv][0] = ich;
      s[wv][1] = icd;
      s[wv][2] = ieh;
      s[wv][3] = ied;
    }
    __syncthreads();
    unsigned long long idx[2] = {tb[0] + ich - ch, tb[1] + icd - cd};
    unsigned long long eb[2] = {tb[2] + ieh - eh, tb[3] + ied - ed};
    for (uint32_t w = 0; w < wv; w++) {
      idx[0] += s[w][0];
      idx[1] += s[w][1];
      eb[0] += s[w][2];
      eb[1] += s[w][3];
    }
    // partitions touched: a plain store of this call's stamp (every writer
    // writes the same value); k_touched counts the stamps afterwards
    if (sel && pf == pl) a.ptouch[pf] = a.stamp;
    uint32_t prevp = 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t r = 0; r < VPL; r++) {
      if (!((sel >> r) & 1u)) continue;
      const int q = ((dev >> r) & 1u) ? Q_DEV : Q_HOST;
      const unsigned long long i = idx[q]++, e0 = eb[q];
      eb[q] += d[r];
      a.qv[q][i] = (uint32_t)(v0 + r);
      a.qe[q][i] = e0;
      if (!a.no_chunks) {
        uint32_t *qc = a.qc[q];
        for (unsigned long long c = (e0 + a.C - 1) / a.C; c * a.C < e0 + d[r]; c++)
          qc[c] = (uint32_t)i;
      }
      if (pf != pl) {
        const uint32_t p = a.vpart[v0 + r];
        if (p != prevp) {
          a.ptouch[p] = a.stamp;
          prevp = p;
        }
      }
    }
    __syncthreads();
  }
}

struct ProbeArgs {
  const uint32_t *front, *rbit, *deg, *hint;
  const uint64_t *eptr;
  uint32_t *vis, *next;
  uint64_t n;
  uint32_t ntiles, D, rec;
  uint32_t *tcnt[2];
  unsigned long long *tdeg[2];
  unsigned long long *ssum[2];
  unsigned long long *cnt;  // [0] hint hits, [1] unvisited, [2] with a hint, [3] words
                            // written, [4] candidates written, [5] candidates read
  uint32_t *slist;          // short lists: wave (tile, w) writes from slot tile*TILE + w*SUB
  uint32_t *scnt;           //   and their count at [tile * 4 + w]
  uint32_t *llist, *lcnt;   // long lists (degree > D), same layout
  // next-frontier bookkeeping (nctr != nullptr): every vertex this level
  // places gets its value here, and nctr sums the entries and edges of the
  // next frontier per residency, so a following bottom-up level needs no
  // count pass
  uint32_t *val;
  uint32_t level_value;
  LevelCtr *nctr;
  // candidates: every vertex still unvisited after the hint test goes to
  // cout (per wave region, in vertex order; count at ccout[region]).  A
  // later probe of the same traversal reads that list (cin/ccin) instead of
  // sweeping every vertex: the unvisited set only shrinks.
  const uint32_t *cin, *ccin;
  uint32_t *cout, *ccout;
  LevelSide side;  // zero words only
};

constexpr uint32_t SHORT_MAX = 16;  // longest list the probe scans in full
constexpr uint32_t PASS = ROUNDS / 4;  // the sweep's rounds per pass

// The probes' counters: one atomic per counter and block (the grid is a
// few blocks per CU).
__device__ __forceinline__ void probe_flush(const ProbeArgs &a, unsigned long long (*s)[10],
                                            uint32_t wv, uint32_t lane, unsigned long long hits,
                                            unsigned long long nun, unsigned long long nhint,
                                            unsigned long long nwords, const uint32_t fn[2],
                                            const unsigned long long fe[2],
                                            unsigned long long ncout, unsigned long long ncin) {
  const unsigned long long c10[10] = {wave_sum(hits),  wave_sum(nun),
                                      wave_sum(nhint), wave_sum(nwords),
                                      wave_sum((unsigned long long)fn[0]),
                                      wave_sum((unsigned long long)fn[1]),
                                      wave_sum(fe[0]), wave_sum(fe[1]),
                                      wave_sum(ncout), wave_sum(ncin)};
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 10; i++) s[wv][i] = c10[i];
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    unsigned long long t = 0;
    for (int w = 0; w < CBS / 64; w++) t += s[w][threadIdx.x];
    if (t) {
      const uint32_t i = threadIdx.x;
      if (i < 4) atomicAdd(&a.cnt[i], t);
      else if (i < 6) atomicAdd(&a.nctr->qn[i - 4], t);
      else if (i < 8) atomicAdd(&a.nctr->qw[i - 6], t);
      else atomicAdd(&a.cnt[i - 4], t);
      if (i == 8 && a.nctr) atomicAdd(&a.nctr->cand, t);
    }
  }
}

// Bottom-up probe, one wave per 1024-vertex sub-tile:
//   1. visited, residency and next-frontier words (lanes 0-31, one load
//      each);
//   2. per pass of PASS rounds (64 vertices a round), hints and degrees
//      loaded together, then the frontier words the unvisited vertices'
//      hints name; a hit places the vertex (it is at level L+1 whichever
//      way the level runs) and is written into the wave's own visited /
//      next words by the lane that loaded them.  Passes keep a pass's
//      loads in flight together within four waves' registers per SIMD
//      (all 16 rounds at once spilled and ran no faster);
//   3. the degrees of the rest.  Lists of 1..D records go to the wave's slots
//      of the short-list array, in vertex order, for k_bu_short; longer
//      ones to the long-list array, for the early-exit kernel; every
//      unvisited vertex without a hint hit to the candidate list, which a
//      later probe of the traversal reads instead of sweeping
//      (k_bu_probe_list).
//   Counts: tile set 0 holds the short lists (entries, records), tile set
//   1 the long unvisited lists (degree > D) left for the early-exit kernel,
//   both split by residency.
__global__ void __launch_bounds__(CBS, 4) k_bu_probe(ProbeArgs a) {
  __shared__ unsigned long long s[CBS / 64][10];
  const uint32_t lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // per-lane counts (each at most n, which fits 32 bits; edge sums take 64)
  uint32_t hits = 0, nun = 0, nhint = 0, nwords = 0, ncout = 0;
  unsigned long long fe[2] = {0, 0};  // next frontier: edges (host, HBM)
  uint32_t fn[2] = {0, 0};            //   and entries
  side_clear(a.side);
  for (uint32_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const uint64_t base = (uint64_t)tile * TILE + (uint64_t)wv * SUB, w0 = base >> 5;
    // the sub-tile's arrays from wave-uniform bases, indexed by 32-bit
    // offsets (scalar base + vector offset addressing); nin: its vertices
    // below n
    const uint32_t nin = base < a.n ? (uint32_t)min<uint64_t>(a.n - base, SUB) : 0u;
    uint32_t *const visb = a.vis + w0;
    uint32_t *const nextb = a.next + w0;
    const uint32_t *const rbitb = a.rbit + w0;
    const uint32_t *const degb = a.deg + base;
    const bool owner = lane < 32 && lane * 32 < nin;
    const uint32_t vw = owner ? visb[lane] : 0xFFFFFFFFu;
    const uint32_t rw = owner ? rbitb[lane] : 0u;
    const uint32_t nw = owner ? nextb[lane] : 0u;  // this wave's own next words
    // fast path: no vertex of the block's tile is unvisited
    uint32_t uw = 0;
    if (owner) {
      uw = ~vw;
      if (nin - lane * 32 < 32) uw &= (1u << (nin - lane * 32)) - 1u;
    }
    if (!__syncthreads_or(uw != 0)) {
      if (lane == 0) {
        a.scnt[tile * (CBS / 64) + wv] = 0;
        a.lcnt[tile * (CBS / 64) + wv] = 0;
        if (a.ccout) a.ccout[tile * (CBS / 64) + wv] = 0;
      }
      if (threadIdx.x < 8) {
        const uint32_t set = threadIdx.x >> 2, i = threadIdx.x & 3;
        if (i < 2) a.tcnt[set][2 * tile + i] = 0;
        else a.tdeg[set][2 * tile + i - 2] = 0;
      }
      continue;
    }
    uint32_t un = 0, dev = 0;  // bit r: round r's vertex is unvisited / HBM-resident
#pragma unroll
    for (uint32_t r = 0; r < ROUNDS; r++) {
      const int src = (int)(2 * r + (lane >> 5));
      const uint32_t bit = 1u << (lane & 31);
      const bool in = r * 64 + lane < nin;
      if (in && !((uint32_t)__shfl((int)vw, src, 64) & bit)) un |= 1u << r;
      if ((uint32_t)__shfl((int)rw, src, 64) & bit) dev |= 1u << r;
    }
    nun += __popc(un);
    uint32_t m = 0;  // hits of word `lane` (lanes 0-31)
    // counts, and the wave's short lists (in vertex order).  With hints, a
    // degree-1 vertex is settled already: its hint is its one neighbour.
    const uint32_t dlo = a.hint ? 2u : 1u;
    // entries: short host/HBM, long host/HBM (wave-uniform); the short
    // lists' records, host/HBM, per lane (at most 16 per vertex).  The long
    // lists' records are not summed: nothing reads them.
    uint32_t xc[4] = {0, 0, 0, 0}, es[2] = {0, 0};
    uint32_t ns = 0, nl = 0, nc = 0;
    // a pass loads its hints and degrees unconditionally (the lines are
    // fetched whole anyway, and both arrays are padded to whole tiles), so
    // that no load waits on a branch
#pragma unroll 1
    for (uint32_t r0 = 0; r0 < ROUNDS; r0 += PASS) {
      uint32_t h[PASS], d[PASS];
      if (a.hint) {
        const uint32_t *const hintb = a.hint + base + r0 * 64;
#pragma unroll
        for (uint32_t r = 0; r < PASS; r++) h[r] = hintb[r * 64 + lane];
      }
      const uint32_t *const degh = degb + r0 * 64;
#pragma unroll
      for (uint32_t r = 0; r < PASS; r++) d[r] = degh[r * 64 + lane];
      uint32_t f = 0;  // bit r: round r0 + r's vertex is a hint hit
      if (a.hint) {
#pragma unroll
        for (uint32_t r = 0; r < PASS; r++) {
          if (!((un >> (r0 + r)) & 1u)) h[r] = NO_HINT;
          nhint += h[r] != NO_HINT;
        }
        // the frontier words the hints name, loaded together (a vertex
        // without a hint reads word 0 and is ignored)
        uint32_t fw[PASS];
#pragma unroll
        for (uint32_t r = 0; r < PASS; r++) fw[r] = a.front[h[r] != NO_HINT ? h[r] >> 5 : 0];
#pragma unroll
        for (uint32_t r = 0; r < PASS; r++)
          if (h[r] != NO_HINT && ((fw[r] >> (h[r] & 31)) & 1u)) f |= 1u << r;
#pragma unroll
        for (uint32_t r = 0; r < PASS; r++) {
          const unsigned long long b = wave_ballot((f >> r) & 1u);
          if (lane == 2 * (r0 + r)) m = (uint32_t)b;
          if (lane == 2 * (r0 + r) + 1) m = (uint32_t)(b >> 32);
        }
        // with the bookkeeping on, the hits keep their degrees for the sums
        if (!a.nctr) un &= ~(f << r0);
      }
#pragma unroll
      for (uint32_t r = 0; r < PASS; r++)
        if (!((un >> (r0 + r)) & 1u)) d[r] = 0;
      if (a.nctr && f) {
#pragma unroll
        for (uint32_t r = 0; r < PASS; r++) {
          if (!((f >> r) & 1u)) continue;
          const uint32_t q = (dev >> (r0 + r)) & 1u;
          (a.val + base + r0 * 64)[r * 64 + lane] = a.level_value;
          fn[q] += 1;
          fe[q] += d[r];
          d[r] = 0;  // placed: no list to scan
        }
        un &= ~(f << r0);
      }
      // the lists in vertex order; the class counts from the ballots (wave-
      // uniform, on the scalar unit), the short lists' records per lane
#pragma unroll
      for (uint32_t r = 0; r < PASS; r++) {
        const uint32_t rr = r0 + r;
        const bool q = (dev >> rr) & 1u;
        const bool sh = d[r] >= dlo && d[r] <= a.D;
        const bool lg = d[r] >= dlo && d[r] > a.D;
        const unsigned long long b = wave_ballot(sh);  // every lane takes part
        const unsigned long long bl = wave_ballot(lg);
        const unsigned long long bq = wave_ballot(q);
        const uint32_t v = (uint32_t)base + rr * 64 + lane;
        if (sh) (a.slist + base)[ns + lanes_below(b)] = v;
        if (lg) (a.llist + base)[nl + lanes_below(bl)] = v;
        ns += __popcll(b);
        nl += __popcll(bl);
        if (a.cout) {
          const bool c = (un >> rr) & 1u;
          const unsigned long long bc = wave_ballot(c);
          if (c) (a.cout + base)[nc + lanes_below(bc)] = v;
          nc += __popcll(bc);
        }
        xc[0] += __popcll(b & ~bq);
        xc[1] += __popcll(b & bq);
        xc[2] += __popcll(bl & ~bq);
        xc[3] += __popcll(bl & bq);
        es[0] += sh && !q ? d[r] : 0u;
        es[1] += sh && q ? d[r] : 0u;
      }
    }
    if (lane == 0) {
      a.scnt[tile * (CBS / 64) + wv] = ns;
      a.lcnt[tile * (CBS / 64) + wv] = nl;
      if (a.ccout) {
        a.ccout[tile * (CBS / 64) + wv] = nc;
        ncout += nc;
      }
    }
    hits += __popc(m);  // hint hits
    if (owner && m) {
      visb[lane] = vw | m;
      nextb[lane] = nw | m;
      nwords++;
    }
    // s[w][set * 4 + i]: i = 0, 1 entries (host, HBM), 2, 3 edges
    const uint32_t e0 = wave_sum(es[0]), e1 = wave_sum(es[1]);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 4; k++) s[wv][(k >> 1) * 4 + (k & 1)] = xc[k];
      s[wv][2] = e0;
      s[wv][3] = e1;
      s[wv][6] = s[wv][7] = 0;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
      unsigned long long y = 0;
      for (int w = 0; w < CBS / 64; w++) y += s[w][threadIdx.x];
      const uint32_t set = threadIdx.x >> 2, i = threadIdx.x & 3;
      if (i < 2) a.tcnt[set][2 * tile + i] = (uint32_t)y;
      else a.tdeg[set][2 * tile + i - 2] = y;
      if (y) atomicAdd(&a.ssum[set][(tile / SEG) * 4 + i], y);
    }
    __syncthreads();
  }
  probe_flush(a, s, wv, lane, hits, nun, nhint, nwords, fn, fe, ncout, 0);
}

// Lanes [lo, hi) of a round whose first flat index is t0, for the run of
// flat indices [o, o + c).
__device__ __forceinline__ unsigned long long run_lanes(uint32_t o, uint32_t c, uint32_t t0) {
  const int64_t lo = (int64_t)o - t0, hi = (int64_t)o + c - t0;
  if (hi <= 0 || lo >= 64 || c == 0) return 0ull;
  const uint32_t l = lo > 0 ? (uint32_t)lo : 0u, h = hi < 64 ? (uint32_t)hi : 64u;
  const unsigned long long upto = h == 64 ? ~0ull : (1ull << h) - 1ull;
  return upto & ~((1ull << l) - 1ull);
}

// List-mode probe: one wave per group of LR consecutive regions, over the
// candidates an earlier probe of the traversal wrote for them (vertices
// unvisited after its hint test; the unvisited set only shrinks).  The
// group's candidates are spread over the wave's lanes, 64 per round, so a
// level with few unvisited vertices costs a round or two per wave instead
// of a sweep of every vertex.  Per candidate: still unvisited (its visited
// bit) and its residency, then the sweep's hint test, degree and list
// classification.  Outputs are the sweep's: short lists, long lists and
// new candidates in each region's slots (lane j keeps region j's running
// counts), hit words per region (collected in LDS, written by this wave,
// which owns the group's words), and the segment sums of both sets.  No
// tile tables: after a probe nothing reads them.
constexpr uint32_t LR = 8;  // regions per group
constexpr uint32_t LU = 4;  // rounds of candidates whose loads are in flight together
static_assert(SEG * (CBS / 64) % LR == 0, "a group lies in one segment");
__global__ void __launch_bounds__(CBS) k_bu_probe_list(ProbeArgs a) {
  __shared__ uint32_t hw[CBS / 64][LR * (SUB / 32)];
  __shared__ unsigned long long s[CBS / 64][10];
  constexpr uint32_t VNONE = 0xFFFFFFFFu;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long lt = lanemask_lt();
  const uint32_t nreg = a.ntiles * (CBS / 64), ngroups = (nreg + LR - 1) / LR;
  const uint32_t nwaves = gridDim.x * (CBS / 64);
  const uint32_t dlo = a.hint ? 2u : 1u;
  unsigned long long hits = 0, nun = 0, nhint = 0, nwords = 0, ncin = 0, ncout = 0;
  unsigned long long fe[2] = {0, 0};
  uint32_t fn[2] = {0, 0};
  side_clear(a.side);
  for (uint32_t gi = blockIdx.x * (CBS / 64) + wv; gi < ngroups; gi += nwaves) {
    const uint32_t r0 = gi * LR;
    // lane j < LR: region r0 + j's candidates [oj, oj + cj) in the group
    const bool rl = lane < LR && r0 + lane < nreg;
    const uint32_t cj = rl ? a.ccin[r0 + lane] : 0u;
    const uint32_t incl = wave_incl_scan(cj), oj = incl - cj;
    const uint32_t T = (uint32_t)__shfl((int)incl, 63, 64);
    if (T == 0) {
      if (rl) {
        a.scnt[r0 + lane] = 0;
        a.lcnt[r0 + lane] = 0;
        if (a.ccout) a.ccout[r0 + lane] = 0;
      }
      continue;
    }
    if (lane == 0) ncin += T;
    for (uint32_t i = lane; i < LR * (SUB / 32); i += 64) hw[wv][i] = 0;
    uint32_t ns = 0, nl = 0, nc = 0;          // lane j: region r0 + j's counts so far
    uint32_t xc[4] = {0, 0, 0, 0};            // entries: short host/HBM, long host/HBM
    unsigned long long xe[4] = {0, 0, 0, 0};  // edges, same order
    // LU rounds of 64 candidates at a time: their loads (candidate ids,
    // then visited, residency, hint and degree words, then the frontier
    // words the hints name) are in flight together; the lists are then
    // written round by round, in candidate order
    for (uint32_t tb = 0; tb < T; tb += 64 * LU) {
      uint32_t v[LU], j[LU];
#pragma unroll
      for (uint32_t u = 0; u < LU; u++) {
        const uint32_t t = tb + u * 64 + lane;
        // the lane's region: the last j whose run starts at or before t
        uint32_t jj = 0;
#pragma unroll
        for (uint32_t q = 1; q < LR; q++)
          if ((uint32_t)__shfl((int)oj, q, 64) <= t) jj = q;
        j[u] = jj;
        const uint32_t oJ = (uint32_t)__shfl((int)oj, jj, 64);
        v[u] = t < T ? a.cin[(uint64_t)(r0 + jj) * SUB + (t - oJ)] : VNONE;
      }
      // a lane past the candidates reads the words of vertex 0 (always a
      // valid vertex) and ignores them
      uint32_t vw[LU], rw[LU], hr[LU], dr[LU];
#pragma unroll
      for (uint32_t u = 0; u < LU; u++) {
        const uint32_t x = v[u] != VNONE ? v[u] : 0u;
        vw[u] = a.vis[x >> 5];
        rw[u] = a.rbit[x >> 5];
        hr[u] = a.hint ? a.hint[x] : NO_HINT;
        dr[u] = a.deg[x];
      }
      uint32_t fw[LU];
#pragma unroll
      for (uint32_t u = 0; u < LU; u++) {
        if (v[u] == VNONE) vw[u] = 0xFFFFFFFFu;
        const bool un = !(vw[u] & (1u << (v[u] & 31)));
        if (!un) hr[u] = NO_HINT;
        fw[u] = a.front[hr[u] != NO_HINT ? hr[u] >> 5 : 0];
      }
#pragma unroll
      for (uint32_t u = 0; u < LU; u++) {
        const uint32_t t0 = tb + u * 64;
        if (t0 >= T) break;  // wave-uniform
        const uint32_t vv = v[u], jj = j[u];
        const uint32_t oJ = (uint32_t)__shfl((int)oj, jj, 64);
        const uint32_t cJ = (uint32_t)__shfl((int)cj, jj, 64);
        const uint64_t rbase = (uint64_t)(r0 + jj) * SUB;
        const uint32_t bit = 1u << (vv & 31);
        bool un = !(vw[u] & bit);
        const uint32_t q = (rw[u] & bit) ? 1u : 0u;
        nun += un;
        const uint32_t h = hr[u];
        nhint += h != NO_HINT;
        const bool hit = h != NO_HINT && ((fw[u] >> (h & 31)) & 1u);
        if (hit) {
          const uint32_t o = (uint32_t)(vv - rbase);
          atomicOr(&hw[wv][jj * (SUB / 32) + (o >> 5)], bit);
        }
        // with the bookkeeping on, a hit's degree is the next frontier's
        uint32_t d = un && (!hit || a.nctr) ? dr[u] : 0u;
        if (hit) {
          if (a.nctr) {
            a.val[vv] = a.level_value;
            fn[q] += 1;
            fe[q] += d;
          }
          d = 0;
          un = false;
        }
        const bool sh = un && d >= dlo && d <= a.D;
        const bool lg = un && d >= dlo && d > a.D;
        const unsigned long long mj = run_lanes(oJ, cJ, t0);
        const unsigned long long bs = __ballot(sh), bl = __ballot(lg), bc = __ballot(un);
        const uint32_t ps = (uint32_t)__shfl((int)ns, jj, 64), pl = (uint32_t)__shfl((int)nl, jj, 64);
        const uint32_t pc = (uint32_t)__shfl((int)nc, jj, 64);
        if (sh) a.slist[rbase + ps + __popcll(bs & lt & mj)] = vv;
        if (lg) a.llist[rbase + pl + __popcll(bl & lt & mj)] = vv;
        if (un && a.cout) a.cout[rbase + pc + __popcll(bc & lt & mj)] = vv;
        if (lane < LR) {  // this round's lanes of region r0 + lane
          const unsigned long long mo = run_lanes(oj, cj, t0);
          ns += __popcll(bs & mo);
          nl += __popcll(bl & mo);
          nc += __popcll(bc & mo);
        }
        if (un && d >= dlo) {
          const uint32_t k = (sh ? 0 : 2) + q;
          xc[k] += 1;
          xe[k] += d;
        }
      }
    }
    if (rl) {
      a.scnt[r0 + lane] = ns;
      a.lcnt[r0 + lane] = nl;
      if (a.ccout) {
        a.ccout[r0 + lane] = nc;
        ncout += nc;
      }
    }
    // the group's hit words (its regions are consecutive, so are they)
    const uint64_t w0 = (uint64_t)r0 * (SUB / 32);
    for (uint32_t i = lane; i < LR * (SUB / 32); i += 64) {
      const uint32_t m = hw[wv][i];
      if (m) {
        a.vis[w0 + i] |= m;
        a.next[w0 + i] |= m;
        nwords++;
        hits += __popc(m);
      }
    }
    // the group's segment sums: s[set * 4 + i], i = 0, 1 entries (host,
    // HBM), 2, 3 edges
    const uint32_t seg = (r0 / (CBS / 64)) / SEG;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const unsigned long long c = wave_sum((unsigned long long)xc[k]);
      const unsigned long long e = wave_sum(xe[k]);
      if (lane == 0) {
        const uint32_t set = (uint32_t)k >> 1, i = (uint32_t)k & 1u;
        if (c) atomicAdd(&a.ssum[set][seg * 4 + i], c);
        if (e) atomicAdd(&a.ssum[set][seg * 4 + 2 + i], e);
      }
    }
  }
  probe_flush(a, s, wv, lane, hits, nun, nhint, nwords, fn, fe, ncout, ncin);
}

struct ShortArgs {
  const uint32_t *slist, *scnt, *deg, *front;
  const uint64_t *eptr;
  uint32_t *vis, *next;
  uint32_t nregions;
  const uint32_t *rbit;     // next-frontier bookkeeping, as in ProbeArgs
  uint32_t *val;
  uint32_t level_value;
  LevelCtr *nctr;
  const uint32_t *fsum;     // frontier summary, nullable
};

// The short lists the probe collected: one wave per region (a wave's slots
// of one probe sub-tile), one list per lane -- its address, its records and
// their frontier words in three rounds of loads (with a frontier summary,
// its words first, and a frontier word only where the summary bit is set).
// A list with a frontier vertex places its vertex.
// The first word of the record at a store address (HBM or mapped host
// memory) as a global-memory load: a generic (flat) load also counts
// against the LDS counter, whose waits drain every load in flight.
__device__ __forceinline__ uint32_t ld_record(unsigned long long addr) {
  return *reinterpret_cast<const __attribute__((address_space(1))) uint32_t *>(addr);
}

template <int REC, bool SUM>
__global__ void __launch_bounds__(CBS) k_bu_short(ShortArgs a) {
  __shared__ unsigned long long sh[CBS / 64];
  const uint32_t lane = threadIdx.x & 63;
  unsigned long long fe[2] = {0, 0}, fn[2] = {0, 0};
  const uint32_t nwaves = gridDim.x * (CBS / 64);
  for (uint32_t rg = blockIdx.x * (CBS / 64) + (threadIdx.x >> 6); rg < a.nregions;
       rg += nwaves) {
    const uint32_t cnt = a.scnt[rg];
    const uint64_t base = (uint64_t)rg * SUB;
    // the lane's next list id loads one step ahead (the last id again at
    // the end of the region)
    uint32_t vn = cnt ? a.slist[base + min(lane, cnt - 1)] : 0u;
    for (uint32_t i = lane; i < cnt; i += 64) {
      const uint32_t v = vn;
      vn = a.slist[base + min(i + 64, cnt - 1)];
      const uint32_t dl = a.deg[v];  // 1..16
      const unsigned long long eb = a.eptr[v];
      // the first EAGER records load unconditionally (the index clamped to
      // the list), so they are all in flight together; most lists are
      // that short
      constexpr uint32_t EAGER = 4;
      uint32_t w[SHORT_MAX];
#pragma unroll
      for (uint32_t k = 0; k < EAGER; k++) {
        const uint32_t x = ld_record(eb + (unsigned long long)min(k, dl - 1) * REC);
        w[k] = k < dl ? x : 0xFFFFFFFFu;
      }
#pragma unroll
      for (uint32_t k = EAGER; k < SHORT_MAX; k++)
        w[k] = k < dl ? ld_record(eb + (unsigned long long)k * REC) : 0xFFFFFFFFu;
      // their frontier words, loaded together (an absent record reads word
      // 0 and is ignored)
      uint32_t fw[SHORT_MAX];
      if (SUM) {
        uint32_t sw[SHORT_MAX];
#pragma unroll
        for (uint32_t k = 0; k < SHORT_MAX; k++)
          sw[k] = w[k] != 0xFFFFFFFFu ? a.fsum[w[k] >> (FSUM_SHIFT + 5)] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < SHORT_MAX; k++) {
          fw[k] = 0u;
          if ((sw[k] >> ((w[k] >> FSUM_SHIFT) & 31)) & 1u) fw[k] = a.front[w[k] >> 5];
        }
      } else {
#pragma unroll
        for (uint32_t k = 0; k < SHORT_MAX; k++)
          fw[k] = a.front[w[k] != 0xFFFFFFFFu ? w[k] >> 5 : 0];
      }
      bool hit = false;
#pragma unroll
      for (uint32_t k = 0; k < SHORT_MAX; k++)
        if (w[k] != 0xFFFFFFFFu && ((fw[k] >> (w[k] & 31)) & 1u)) hit = true;
      if (hit) {
        atomicOr(&a.vis[v >> 5], 1u << (v & 31));
        atomicOr(&a.next[v >> 5], 1u << (v & 31));
        if (a.nctr) {
          const uint32_t q = (a.rbit[v >> 5] >> (v & 31)) & 1u;
          a.val[v] = a.level_value;
          fn[q] += 1;
          fe[q] += dl;
        }
      }
    }
  }
  if (a.nctr) {  // every thread takes part in the block sums
    const unsigned long long x0 = block_sum<CBS>(fn[0], sh), x1 = block_sum<CBS>(fn[1], sh);
    const unsigned long long y0 = block_sum<CBS>(fe[0], sh), y1 = block_sum<CBS>(fe[1], sh);
    if (threadIdx.x == 0) {
      if (x0) atomicAdd(&a.nctr->qn[0], x0);
      if (x1) atomicAdd(&a.nctr->qn[1], x1);
      if (y0) atomicAdd(&a.nctr->qw[0], y0);
      if (y1) atomicAdd(&a.nctr->qw[1], y1);
    }
  }
}

// Bases of the probe's long-list regions: one block per segment of SEG
// tiles (SEG * 4 regions, one per thread), on top of the segment's base
// from k_sscan (both residencies).  The last block writes the total after
// the last region, so the early-exit kernel can split the lists evenly.
// Block 0 also publishes the probe's counters (side), all final by now.
struct LbaseArgs {
  const uint32_t *lcnt;
  const unsigned long long *sbase;  // set 1 (long lists) segment bases
  uint32_t *loff;
  uint32_t nregions, nseg;
  LevelSide side;
};
constexpr int LBS = SEG * (CBS / 64);
__global__ void __launch_bounds__(LBS) k_lbase(LbaseArgs a) {
  __shared__ uint32_t sh[LBS / 64 + 1];
  const uint32_t rg = blockIdx.x * LBS + threadIdx.x;
  const uint32_t x = rg < a.nregions ? a.lcnt[rg] : 0u;
  uint32_t tot;
  const uint32_t ex = block_excl_scan<LBS>(x, sh, &tot);
  const uint32_t base = (uint32_t)(a.sbase[blockIdx.x * 4] + a.sbase[blockIdx.x * 4 + 1]);
  if (rg < a.nregions) a.loff[rg] = base + ex;
  if (blockIdx.x == a.nseg - 1 && threadIdx.x == 0) a.loff[a.nregions] = base + tot;
  if (blockIdx.x == 0 && threadIdx.x < 64) side_publish(a.side, threadIdx.x);
}

// Partitions stamped by this call, per residency.
__global__ void __launch_bounds__(CBS) k_touched(CompactArgs a) {
  __shared__ unsigned long long sh[CBS / 64];
  uint32_t c[2] = {0, 0};
  for (uint32_t p = blockIdx.x * CBS + threadIdx.x; p < a.P; p += gridDim.x * CBS)
    if (a.ptouch[p] == a.stamp) c[a.pres[p] ? Q_DEV : Q_HOST]++;
  const unsigned long long h = block_sum<CBS>((unsigned long long)c[0], sh);
  const unsigned long long d = block_sum<CBS>((unsigned long long)c[1], sh);
  if (threadIdx.x == 0) {
    if (h) atomicAdd(&a.ctr->touched[Q_HOST], (unsigned)h);
    if (d) atomicAdd(&a.ctr->touched[Q_DEV], (unsigned)d);
  }
}

// A sparse frontier's queue (launch_queue_sparse).  k_list_front: one pass
// over the bitmap, four words per thread, appends the frontier vertices to
// `list` (one atomic per wave that holds any; the caller bounds the
// frontier, so few waves do) and clears side.zero as side work.
__global__ void __launch_bounds__(CBS) k_list_front(const uint32_t *front, uint64_t words,
                                                    uint32_t *list, uint32_t cap,
                                                    unsigned int *count, LevelSide side) {
  side_clear(side);
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t b = (uint64_t)blockIdx.x * CBS; b * 4 < words; b += (uint64_t)gridDim.x * CBS) {
    const uint64_t i = b + threadIdx.x;  // words [4i, 4i + 4)
    uint32_t x[4] = {0u, 0u, 0u, 0u};
    if (4 * i + 3 < words) {
      const uint4 q = reinterpret_cast<const uint4 *>(front)[i];
      x[0] = q.x;
      x[1] = q.y;
      x[2] = q.z;
      x[3] = q.w;
    } else {
#pragma unroll
      for (uint32_t k = 0; k < 4; k++)
        if (4 * i + k < words) x[k] = front[4 * i + k];
    }
    const uint32_t c = __popc(x[0]) + __popc(x[1]) + __popc(x[2]) + __popc(x[3]);
    const uint32_t inc = wave_incl_scan(c);
    const uint32_t tot = (uint32_t)__shfl((int)inc, 63, 64);
    if (tot == 0) continue;  // wave-uniform
    uint32_t o = 0;
    if (lane == 0) o = atomicAdd(count, tot);
    o = (uint32_t)__shfl((int)o, 0, 64) + inc - c;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      for (uint32_t wd = x[k]; wd; wd &= wd - 1, o++)
        if (o < cap) list[o] = (uint32_t)((4 * i + k) * 32 + (uint32_t)(__ffs(wd) - 1));
    }
  }
}

// k_queue_list: one block writes the queue of the listed vertices (with
// dmin <= degree <= dmax), in list order: entries and edge offsets per
// residency from block scans, the chunk starts, and the totals with the
// partitions touched (the first to stamp a partition counts it); ctr's cand
// is kept and the list count reset.
constexpr int QBS = 1024;
__global__ void __launch_bounds__(QBS) k_queue_list(CompactArgs a, const uint32_t *list,
                                                    uint32_t cap) {
  __shared__ unsigned long long sh[6][QBS / 64 + 1];
  unsigned int *count = &a.ctr->pad[0];
  const uint32_t cnt = min(*count, cap);
  const uint32_t per = (cnt + QBS - 1) / QBS;
  const uint32_t i0 = min(threadIdx.x * per, cnt), i1 = min(i0 + per, cnt);
  unsigned long long n[2] = {0, 0}, e[2] = {0, 0};
  for (uint32_t i = i0; i < i1; i++) {
    const uint32_t v = list[i], d = a.deg[v];
    if (d < a.dmin || d > a.dmax) continue;
    const int q = (a.rbit[v >> 5] >> (v & 31)) & 1u ? Q_DEV : Q_HOST;
    n[q] += 1;
    e[q] += d;
  }
  unsigned long long tn[2], te[2], idx[2], eb[2];
  idx[0] = block_excl_scan<QBS>(n[0], sh[0], &tn[0]);
  idx[1] = block_excl_scan<QBS>(n[1], sh[1], &tn[1]);
  eb[0] = block_excl_scan<QBS>(e[0], sh[2], &te[0]);
  eb[1] = block_excl_scan<QBS>(e[1], sh[3], &te[1]);
  unsigned long long th[2] = {0, 0};
  for (uint32_t i = i0; i < i1; i++) {
    const uint32_t v = list[i], d = a.deg[v];
    if (d < a.dmin || d > a.dmax) continue;
    const int q = (a.rbit[v >> 5] >> (v & 31)) & 1u ? Q_DEV : Q_HOST;
    const unsigned long long k = idx[q]++, e0 = eb[q];
    eb[q] += d;
    a.qv[q][k] = v;
    a.qe[q][k] = e0;
    for (unsigned long long c = (e0 + a.C - 1) / a.C; c * a.C < e0 + d; c++)
      a.qc[q][c] = (uint32_t)k;
    const uint32_t p = a.vpart[v];
    if (atomicExch(&a.ptouch[p], a.stamp) != a.stamp) th[a.pres[p] ? Q_DEV : Q_HOST]++;
  }
  const unsigned long long t0 = block_sum<QBS>(th[0], sh[4]);
  const unsigned long long t1 = block_sum<QBS>(th[1], sh[5]);
  if (threadIdx.x == 0) {
    LevelCtr t = *a.ctr;  // keeps cand
    t.qn[0] = tn[0];
    t.qn[1] = tn[1];
    t.qw[0] = te[0];
    t.qw[1] = te[1];
    t.touched[0] = (unsigned)t0;
    t.touched[1] = (unsigned)t1;
    t.rd[0] = t.rd[1] = 0;
    t.pad[0] = 0;  // the list count, for the next sparse level
    *a.ctr = t;
  }
}

// The frontier summary (launch_fsum): one lane per summary bit, over its 8
// frontier words; lanes 0 and 32 write the wave's two summary words (the
// bits past the last group read as zero).
__global__ void __launch_bounds__(CBS) k_fsum(const uint32_t *front, uint64_t words,
                                             uint32_t *fsum, uint64_t nbits) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t b = blockIdx.x * (uint64_t)CBS + threadIdx.x; b - lane < nbits;
       b += (uint64_t)gridDim.x * CBS) {
    uint32_t any = 0;
    if (b < nbits) {
#pragma unroll
      for (uint32_t k = 0; k < 8; k++)
        if (b * 8 + k < words) any |= front[b * 8 + k];
    }
    const unsigned long long m = __ballot(any != 0);
    if (lane == 0) fsum[b >> 5] = (uint32_t)m;
    if (lane == 32) fsum[b >> 5] = (uint32_t)(m >> 32);
  }
}

}  // namespace

void launch_fsum(Graph &g, const uint32_t *front, uint32_t *fsum, hipStream_t s) {
  const uint64_t nbits = fsum_bits(g.n);
  const uint32_t grid = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(div_up(nbits, CBS), 8ull * g.cus));
  hipLaunchKernelGGL(k_fsum, dim3(grid), dim3(CBS), 0, s, front, bitmap_words(g.n), fsum, nbits);
  HIPC(hipGetLastError());
}

void launch_queue_sparse(Graph &g, QueueSet &w, const uint32_t *front, LevelCtr *ctr,
                         hipStream_t s, const LevelSide *side) {
  if (side && (side->pub || side->xpub || side->zctr))
    fail(THV_ESTATE, "the sparse queue only clears words");
  CompactArgs a;
  memset(&a, 0, sizeof a);
  a.rbit = g.store.d_rbit.as<uint32_t>();
  a.deg = g.d_deg.as<uint32_t>();
  a.vpart = g.store.d_vpart.as<uint32_t>();
  a.pres = g.store.d_pres.as<uint8_t>();
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
  a.dmin = 1;
  a.dmax = 0xFFFFFFFFu;
  LevelSide zs;
  if (side) zs = *side;
  const uint64_t words = bitmap_words(g.n);
  const uint32_t grid = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(div_up(words, 4 * CBS), 4ull * g.cus));
  hipLaunchKernelGGL(k_list_front, dim3(grid), dim3(CBS), 0, s, front, words,
                     w.slist.as<uint32_t>(), SPARSE_MAX, &ctr->pad[0], zs);
  hipLaunchKernelGGL(k_queue_list, dim3(1), dim3(QBS), 0, s, a, w.slist.as<const uint32_t>(),
                     SPARSE_MAX);
  HIPC(hipGetLastError());
}

void launch_compact(Graph &g, QueueSet &w, const uint32_t *front, LevelCtr *ctr,
                    hipStream_t s, int mode, uint32_t *val, uint32_t val_value, uint32_t dmin,
                    uint32_t dmax, int tset, const uint32_t *fval, uint32_t fthr,
                    uint32_t *clear, uint32_t *minout, const LevelSide *side) {
  if (dmin == 0) dmin = 1;  // degree-0 vertices never have work
  Store &st = g.store;
  CompactArgs a;
  memset(&a, 0, sizeof a);
  a.front = front;
  a.rbit = st.d_rbit.as<uint32_t>();
  a.deg = g.d_deg.as<uint32_t>();
  a.vpart = st.d_vpart.as<uint32_t>();
  a.pres = st.d_pres.as<uint8_t>();
  a.P = st.P;
  a.n = g.n;
  a.ntiles = (uint32_t)std::max<uint64_t>(1, div_up(g.n, TILE));
  a.tcnt = w.tcnt[tset].as<uint32_t>();
  a.tdeg = w.tdeg[tset].as<unsigned long long>();
  a.ssum = w.ssum[tset].as<unsigned long long>();
  a.sbase = w.sbase[tset].as<unsigned long long>();
  a.nseg = (uint32_t)div_up(a.ntiles, SEG);
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
  a.val = (mode & (CMP_INVERT | CMP_WRITE_ONLY)) ? nullptr : val;
  a.val_value = val_value;
  a.dmin = dmin;
  a.dmax = dmax;
  a.fval = fval;
  a.fthr = fthr;
  a.clear = clear;
  a.minout = minout;
  if (side) a.side = *side;
  const uint32_t grid = a.ntiles < 65535u ? a.ntiles : 65535u;
  // WRITE_ONLY continues a COUNT_ONLY call on the same bitmap and counters:
  // the tile bases and totals are still in place
  if (!(mode & CMP_WRITE_ONLY)) {
    hipLaunchKernelGGL(k_tcount, dim3(grid), dim3(CBS), 0, s, a);
    hipLaunchKernelGGL(k_sscan, dim3(1), dim3(SBS), 0, s, a);
  }
  if (!(mode & CMP_COUNT_ONLY)) {
    hipLaunchKernelGGL(k_twrite, dim3(grid), dim3(CBS), 0, s, a);
    const uint32_t tg = (uint32_t)std::min<uint64_t>(div_up(st.P, CBS), 64);
    hipLaunchKernelGGL(k_touched, dim3(tg ? tg : 1), dim3(CBS), 0, s, a);
  }
  HIPC(hipGetLastError());
}

void launch_bu_probe(Graph &g, QueueSet &w, const uint32_t *front, uint32_t *vis,
                     uint32_t *next, const uint32_t *hint, uint32_t D, LevelCtr *cs,
                     LevelCtr *cl, unsigned long long *cnt, hipStream_t s, hipEvent_t t0,
                     hipEvent_t t1, LevelCtr *nctr, uint32_t *val, uint32_t level_value,
                     const uint32_t *cin, const uint32_t *ccin, uint32_t *cout,
                     uint32_t *ccout, const LevelSide *side, const uint32_t *fsum) {
  ProbeArgs p;
  memset(&p, 0, sizeof p);
  if (side && side->zctr) fail(THV_ESTATE, "the probe resets no counter");
  if (side) {
    p.side.zero = side->zero;
    p.side.nzero = side->nzero;
  }
  p.front = front;
  p.rbit = g.store.d_rbit.as<uint32_t>();
  p.deg = g.d_deg.as<uint32_t>();
  p.hint = hint;
  p.vis = vis;
  p.next = next;
  p.n = g.n;
  p.ntiles = (uint32_t)std::max<uint64_t>(1, div_up(g.n, TILE));
  p.D = D;
  for (int k = 0; k < 2; k++) {
    p.tcnt[k] = w.tcnt[k].as<uint32_t>();
    p.tdeg[k] = w.tdeg[k].as<unsigned long long>();
    p.ssum[k] = w.ssum[k].as<unsigned long long>();
  }
  p.cnt = cnt;
  p.eptr = g.store.d_eptr.as<uint64_t>();
  p.rec = (uint32_t)g.rec;
  if (D > SHORT_MAX) fail(THV_EINVAL, "bottom-up full-scan degree above 16");
  p.slist = w.slist.as<uint32_t>();
  p.scnt = w.scnt.as<uint32_t>();
  p.llist = w.llist.as<uint32_t>();
  p.lcnt = w.lcnt.as<uint32_t>();
  p.nctr = nctr;
  p.val = val;
  p.level_value = level_value;
  if (nctr && !val) fail(THV_ESTATE, "next-frontier bookkeeping needs the value array");
  if ((cin == nullptr) != (ccin == nullptr) || (cout == nullptr) != (ccout == nullptr) ||
      (cin && cin == cout))
    fail(THV_ESTATE, "probe candidate lists: both arrays, and distinct input and output");
  p.cin = cin;
  p.ccin = ccin;
  p.cout = cout;
  p.ccout = ccout;
  const uint32_t grid = std::min<uint32_t>(p.ntiles, (uint32_t)(g.shape.probe_bpc * g.cus));
  if (t0) HIPC(hipEventRecord(t0, s));
  if (cin) {
    const uint32_t groups = (uint32_t)div_up((uint64_t)p.ntiles * (CBS / 64), LR);
    const uint32_t lgrid = std::min<uint32_t>((uint32_t)div_up(groups, CBS / 64), 64u * g.cus);
    hipLaunchKernelGGL(k_bu_probe_list, dim3(lgrid), dim3(CBS), 0, s, p);
  } else {
    hipLaunchKernelGGL(k_bu_probe, dim3(grid), dim3(CBS), 0, s, p);
  }
  if (D > 0) {
    ShortArgs sa;
    sa.slist = p.slist;
    sa.scnt = p.scnt;
    sa.deg = p.deg;
    sa.front = front;
    sa.eptr = p.eptr;
    sa.vis = vis;
    sa.next = next;
    sa.nregions = p.ntiles * (CBS / 64);
    sa.rbit = p.rbit;
    sa.val = val;
    sa.level_value = level_value;
    sa.nctr = nctr;
    sa.fsum = fsum;
    if (g.rec == 8) {
      if (fsum) hipLaunchKernelGGL((k_bu_short<8, true>), dim3(grid), dim3(CBS), 0, s, sa);
      else hipLaunchKernelGGL((k_bu_short<8, false>), dim3(grid), dim3(CBS), 0, s, sa);
    } else {
      if (fsum) hipLaunchKernelGGL((k_bu_short<4, true>), dim3(grid), dim3(CBS), 0, s, sa);
      else hipLaunchKernelGGL((k_bu_short<4, false>), dim3(grid), dim3(CBS), 0, s, sa);
    }
  }
  if (t1) HIPC(hipEventRecord(t1, s));
  // tile bases and totals of both sets
  CompactArgs sa2[2];
  for (int k = 0; k < 2; k++) {
    CompactArgs &a = sa2[k];
    memset(&a, 0, sizeof a);
    a.ntiles = p.ntiles;
    a.tcnt = w.tcnt[k].as<uint32_t>();
    a.tdeg = w.tdeg[k].as<unsigned long long>();
    a.ssum = w.ssum[k].as<unsigned long long>();
    a.sbase = w.sbase[k].as<unsigned long long>();
    a.nseg = (uint32_t)div_up(a.ntiles, SEG);
    a.ctr = k == 0 ? cs : cl;
  }
  hipLaunchKernelGGL(k_sscan2, dim3(2), dim3(SBS), 0, s, sa2[0], sa2[1]);
  LbaseArgs lb;
  lb.lcnt = p.lcnt;
  lb.sbase = w.sbase[1].as<unsigned long long>();
  lb.loff = w.loff.as<uint32_t>();
  lb.nregions = p.ntiles * (CBS / 64);
  lb.nseg = (uint32_t)div_up(p.ntiles, SEG);
  if (side) {
    lb.side = *side;
    lb.side.zero = nullptr;
    lb.side.nzero = 0;
  }
  hipLaunchKernelGGL(k_lbase, dim3(lb.nseg), dim3(LBS), 0, s, lb);
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
