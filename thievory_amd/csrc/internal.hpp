// internal.hpp -- shared definitions of the MI355X graph engine.
//
// Layout in HBM (per graph):
//   deg[n]    u32  edge-list length of every vertex (out-degree for a CSR,
//                  in-degree for a CSC)
//   eptr[n]   u64  address of the vertex's first edge record, either in the
//                  device slab (partition resident in HBM) or in the mapped
//                  pinned host copy (read zero-copy over the host link)
//   pv[P+1]   u32  first vertex of every partition; pres[P] u8 residency
//   slab           the HBM-resident partitions' edge records, packed
// plus the per-algorithm workspace (frontier bitmaps, queues, values).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/thv.h"

namespace thv {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string &msg);
void hip_fail(hipError_t e, const char *what, const char *file, int line);

#define HIPC(x)                                                   \
  do {                                                            \
    hipError_t _e = (x);                                          \
    if (_e != hipSuccess) ::thv::hip_fail(_e, #x, __FILE__, __LINE__); \
  } while (0)

// ---- device memory helpers ------------------------------------------------

struct DevMem {
  void *p = nullptr;
  size_t bytes = 0;
  bool own = true;  // false: a view of memory someone else holds (borrow)
  DevMem() = default;
  DevMem(const DevMem &) = delete;
  DevMem &operator=(const DevMem &) = delete;
  ~DevMem() { release(); }
  void alloc(size_t b);
  // points at b bytes the caller holds; release() then forgets them
  void borrow(void *q, size_t b) {
    release();
    p = q;
    bytes = b;
    own = false;
  }
  void release();
  template <class T> T *as() const { return static_cast<T *>(p); }
};

struct HostPinned {  // hipHostMalloc'd, mapped into the GPU address space
  void *p = nullptr;
  void *dev = nullptr;
  size_t bytes = 0;
  HostPinned() = default;
  HostPinned(const HostPinned &) = delete;
  HostPinned &operator=(const HostPinned &) = delete;
  ~HostPinned() { release(); }
  void alloc(size_t b);
  void release();
};

// ---- store -------------------------------------------------------------------

struct Store {
  bool configured = false;
  uint32_t part_edges = 1u << 16;
  uint64_t budget = UINT64_MAX;      // device edge budget in bytes
  uint32_t P = 0;                    // partitions
  uint32_t Pdev = 0;                 // partitions resident in HBM
  uint64_t dev_edges = 0;            // edges resident in HBM
  uint64_t host_vertices = 0;        // vertices with edges in host memory
  uint64_t host_short_lists = 0;     //   of them with 1..16 records
  uint64_t host_short_edges = 0;     //   and those lists' records
  std::vector<uint32_t> pv;          // P+1 first vertex of each partition
  std::vector<uint8_t> pres;         // P residency (1 = HBM)
  DevMem slab;                       // resident records
  DevMem side;                       // short host lists copied to HBM
  uint64_t side_edges = 0;           //   (bottom-up layout; in dev_edges)
  DevMem d_pv, d_pres, d_eptr;
  DevMem d_rbit;                     // n bits: vertex's edges are in HBM
  DevMem d_vpart;                    // n x u32: partition of each vertex
  // Parent hints (symmetric graphs): n x u32, for a vertex with edges its
  // neighbour of largest degree, else NO_HINT.  A split store takes their
  // bytes out of the budget; a resident store holds them beside the lists.
  bool hint_reserved = false;
  DevMem d_hint;
  bool sparse_first = false;         // residency ranked sparsest partitions first
};

constexpr uint32_t NO_HINT = 0xFFFFFFFFu;

// ---- workspace (per graph, allocated on first use) --------------------------

enum { Q_HOST = 0, Q_DEV = 1 };

// Counters written by the compaction kernel for one level.
struct LevelCtr {
  unsigned long long qn[2];   // queue entries (host, device)
  unsigned long long qw[2];   // queue edges
  unsigned int touched[2];    // partitions with >=1 active vertex
  unsigned int pad[2];
  unsigned long long rd[2];   // bottom-up BFS: edge records read (host, HBM)
  unsigned long long cand;    // bottom-up BFS: candidates the level's probe left
};

// [0], [1]: alternating levels; [2]: bottom-up early-exit queue; [3]:
// bottom-up full-scan queue; [4]: next frontier placed by a bottom-up level
enum { CTR_SLOTS = 5 };

struct Graph;

// Frontier queues of one traversal engine: filled by the compaction kernels,
// consumed by an expansion kernel.  Engines that run concurrently on
// different streams own separate QueueSets.
constexpr uint32_t TILE = 4096;  // vertices per compaction tile
constexpr uint32_t PROBE_REGION = 1024;  // vertices per wave region of the bottom-up probe

struct QueueSet {
  uint64_t n = 0, m = 0;
  uint32_t C = 0;             // edges per expansion chunk
  uint32_t P = 0;             // partitions the touch stamps cover
  uint32_t stamp = 0;         // compaction call counter (touch stamps)
  DevMem qv[2], qe[2], qc[2]; // queue vertex ids, edge prefixes, chunk starts
  DevMem ctr;                 // LevelCtr[2]
  // per tile and queue, in two sets (a bottom-up level counts its short
  // and long unvisited lists in one sweep):
  DevMem tcnt[2], tdeg[2];    // active vertices / edges
  DevMem ssum[2], sbase[2];   // per segment of 64 tiles: sums (kept zero
                              // between counts) and exclusive bases
  DevMem ptouch;              // per partition: stamp of the last touching call
  DevMem slist, scnt;         // bottom-up probe: short lists per wave region
  DevMem llist, lcnt;         //   and long lists (for the early-exit kernel)
  DevMem loff;                //   with their regions' exclusive bases (+ total)
  DevMem clist[2], ccnt[2];   //   and candidates for the traversal's next probe
  void ensure(const Graph &g, uint32_t C);
};

struct Workspace {
  uint64_t n = 0, m = 0;
  QueueSet qs;
  // frontier bitmaps (n bits, word padded).  The direction-optimising BFS
  // rotates all three: level L reads front[L % 3], writes front[(L+1) % 3]
  // and one of its own kernels clears front[(L+2) % 3] for level L+1.  The
  // other loops use [0] and [1].
  DevMem front[3];
  DevMem vis;                 // BFS visited bitmap
  DevMem fsum;                // frontier summary (launch_fsum)
  DevMem val;                 // u32 values (BFS / CC / SSSP)
  DevMem f0, f1, f2;          // PR arrays (fp32)
  DevMem aux;                 // misc device counters (16 x u64)
  // pinned, mapped readback (CTR_SLOTS LevelCtr, 16 u64): the host reads
  // them after a stream sync; kernels that finish a level write them
  // directly through the device addresses h_ctr_dev / h_aux_dev
  LevelCtr *h_ctr = nullptr;
  unsigned long long *h_aux = nullptr;
  LevelCtr *h_ctr_dev = nullptr;
  unsigned long long *h_aux_dev = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // [2], [3]: probe
  ~Workspace();
};

// ---- launch shape -------------------------------------------------------------

struct Shape {
  int blocks_per_cu = 4;      // expansion grid = blocks_per_cu * CUs
  int ept = 8;                // edges per thread per chunk (chunk = 256*ept)
  int probe_bpc = 8;          // bottom-up probe, short-list and early-exit
                              // grids = probe_bpc * CUs (persistent)
  int batch_bpc = 8;          // multi-source bottom-up grids = batch_bpc * CUs
};

// ---- graph -------------------------------------------------------------------

struct Graph {
  uint64_t n = 0, m = 0;
  int rec = 4;                // edge record bytes (4 = dst, 8 = {dst, w})
  bool csc = false;           // edge lists are in-lists (PageRank pull)
  std::vector<uint64_t> h_off;  // n+1
  HostPinned h_edges;         // m records, mapped
  DevMem d_deg;               // n u32
  DevMem d_iso;               // bitmap: degree 0 (never reached on a symmetric graph)
  DevMem d_outdeg;            // n u32 (CSC only)
  Store store;
  Workspace ws;
  Shape shape;
  bool tuned = false;
  hipStream_t stream = nullptr;
  int device = 0;
  int cus = 256;
  std::mutex mu;              // serialises algorithm calls on one handle
  std::atomic<int> servers{0};  // running servers (store must stay fixed)
  bool sym = false;           // every edge has its reverse
  bool dir_opt = true;        // direction-optimising BFS when sym
  double do_alpha = 14.0, do_beta = 24.0;
  // link cost model of a bottom-up level over a split store: each unvisited
  // vertex with host-resident edges costs one small zero-copy read
  double do_vertex_ns = 3.0;  // per long unvisited host list (measured on MI355X)
  double do_link_gbps = 55.0; // sustained zero-copy read rate
  // bottom-up levels: unvisited vertices of at most this degree are scanned
  // in full by the edge-balanced kernel, longer lists by the early-exit one
  uint32_t do_pull_degree = 16;
  bool parent_hints = true;
  // SSSP: bucket width of delta-stepping (0: frontier Bellman-Ford)
  uint32_t sssp_delta = 16;
  bool cc_union = true;       // CC of a symmetric graph by union-find
  uint32_t uf_max_steps = 1u << 20;  // union-find step guard (then label propagation)
  bool kernel_timing = true;  // BFS: HIP events around its launches
  // bottom-up levels scan with a frontier summary when frontier vertices x
  // do_summary_ratio < n (0: never)
  double do_summary_ratio = 1024.0;
  // store residency order: 0 chosen by the hints (sparsest first with them),
  // 1 densest partitions first, 2 sparsest first with the short-list side area
  int store_layout = 0;
  // PageRank push: the destinations that receive the most records (from a
  // 1-in-64 sample of the edge array), sorted, and their bitmap; blocks
  // accumulate adds to them in LDS.  Built on the first push call.
  DevMem d_hubs, d_hubbit;
  uint32_t nhubs = 0;
  bool hubs_built = false;
  ~Graph();
};

// graph.cpp
void graph_init_device(Graph &g);
void graph_build_tables(Graph &g);
void store_configure(Graph &g, uint64_t budget, uint32_t part_edges);
void store_ensure(Graph &g);
bool store_wants_hints(const Graph &g, uint64_t budget);
// whether the store would rank partitions sparsest first (with `hints`)
bool store_sparse_first(const Graph &g, bool hints);
// whether the current options would build a different store
bool store_stale(const Graph &g);
void store_refresh(Graph &g);
void workspace_ensure(Graph &g);
uint64_t now_ns();

// kernels (frontier.hip / expand.hip / algos.hip)
// A_HINT: parent-hint build -- key64[v] = max over v's edges (v, w) of
// (deg[w] << 32 | ~w).  A_CC_LINK: union-find link of every edge (val is
// the parent array, err the step-guard flag).
enum Algo {
  A_BFS = 0, A_CC = 1, A_SSSP = 2, A_PR_PUSH = 3, A_PR_PULL = 4, A_BFS_NEXT = 5, A_HINT = 6,
  A_CC_LINK = 7
};

struct ExpandArgs {
  const uint32_t *qv[2];
  const unsigned long long *qe[2];
  const uint32_t *qc[2];
  const LevelCtr *ctr;
  const uint64_t *eptr;
  uint32_t *val;            // BFS/CC/SSSP values
  uint32_t *vis;            // BFS visited bitmap
  uint32_t *next;           // next frontier bitmap
  const uint32_t *deg;      // hint build: degrees
  unsigned long long *key64;  // hint build: per-vertex best key
  uint32_t *err;            // CC link: step-guard flag
  uint32_t uf_steps;        // CC link: step guard
  float *pr_src;            // PR push: delta; PR pull: contribution
  float *pr_dst;            // PR push: residual; PR pull: sum
  const uint32_t *hubs;     // PR push: hub ids (nhubs, sorted) and bitmap
  const uint32_t *hubbit;
  uint32_t nhubs;
  uint32_t level_value;     // BFS: value given to discovered vertices
  uint32_t C;
  int rec;
  // A_BFS_NEXT: BFS that also writes level_value for every vertex it
  // discovers and sums them into nctr (entries and edges per residency, deg
  // and rbit), so the next level needs no count pass
  const uint32_t *rbit;
  LevelCtr *nctr;
};

// Level bookkeeping a launch does on the side, so a level needs no memset
// or readback copy of its own (every field may be null / 0):
//   zero[0, nzero): words the launch's grid clears (a bitmap no kernel of
//     the launch reads or writes otherwise);
//   pub_src -> pub: npub LevelCtr copied, once the launch's counters are
//     final, to the host-mapped readback; xsrc -> xpub: nx u64 counters too;
//   zctr: a LevelCtr reset to zero after that copy.
// The host reads the readback after a stream sync: the launch's completion
// orders its stores to host memory before the sync returns.
struct LevelSide {
  uint32_t *zero = nullptr;
  uint64_t nzero = 0;
  const LevelCtr *pub_src = nullptr;
  LevelCtr *pub = nullptr;
  uint32_t npub = 0;
  const unsigned long long *xsrc = nullptr;
  unsigned long long *xpub = nullptr;
  uint32_t nx = 0;
  LevelCtr *zctr = nullptr;
};

// Compaction modes: INVERT takes the vertices whose bit is CLEAR (the
// unvisited set of a bottom-up BFS level), NO_CHUNKS skips the chunk-start
// table, COUNT_ONLY stops after the totals.
enum { CMP_INVERT = 1, CMP_NO_CHUNKS = 2, CMP_COUNT_ONLY = 4, CMP_WRITE_ONLY = 8 };
// val (BFS only): every vertex with its bit set gets val_value, written in
// the counting pass.
// tset: which tile-table set holds the counts (WRITE_ONLY reads it).
// fval / fthr: select only vertices with fval[v] < fthr.  clear: the write
// pass clears the bits of the vertices it selects there (the front bitmap
// itself, for a pending set).  minout (with fval): the count pass lowers
// *minout to the smallest fval of a selected vertex.  side (count pass
// only; ignored by WRITE_ONLY): the count kernel clears side->zero, the
// scan publishes after writing the totals and resets side->zctr.
void launch_compact(Graph &g, QueueSet &qs, const uint32_t *front, LevelCtr *ctr,
                    hipStream_t s, int mode = 0, uint32_t *val = nullptr,
                    uint32_t val_value = 0, uint32_t dmin = 1,
                    uint32_t dmax = 0xFFFFFFFFu, int tset = 0,
                    const uint32_t *fval = nullptr, uint32_t fthr = 0,
                    uint32_t *clear = nullptr, uint32_t *minout = nullptr,
                    const LevelSide *side = nullptr);
// The queue of a frontier of at most SPARSE_MAX vertices (the caller knows
// its size) without tile tables: one pass over the bitmap lists its
// vertices (into qs.slist, counted in ctr->pad[0], which must be 0 and is
// left 0) and one block writes the queue and ctr's totals as launch_compact
// would, with the touched partitions; side (may be null): words to clear.
constexpr uint32_t SPARSE_MAX = 4096;
void launch_queue_sparse(Graph &g, QueueSet &qs, const uint32_t *front, LevelCtr *ctr,
                         hipStream_t s, const LevelSide *side);
// Bottom-up probe, one sweep over the vertices: every unvisited vertex
// whose parent hint (hint may be null) is in front joins vis and next, and
// so does every unvisited vertex of degree 1..D with a frontier vertex in
// its list (scanned in full).  Tile set 0 / *cs count those short lists
// (entries, records read); tile set 1 / *cl the long unvisited lists
// (degree > D) for a WRITE_ONLY compaction of the inverted vis bitmap.
// cnt[0..3] accumulate hint hits, unvisited vertices, unvisited vertices
// with a hint, and bitmap words written.  Events t0 / t1 (may be null)
// bracket the probe kernel alone.  side: the probe kernel clears
// side->zero; the last launch (long-list bases, after the probe, the short
// lists and both scans) publishes (no zctr here).
void launch_bu_probe(Graph &g, QueueSet &qs, const uint32_t *front, uint32_t *vis,
                     uint32_t *next, const uint32_t *hint, uint32_t D, LevelCtr *cs,
                     LevelCtr *cl, unsigned long long *cnt, hipStream_t s,
                     hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr,
                     LevelCtr *nctr = nullptr, uint32_t *val = nullptr,
                     uint32_t level_value = 0, const uint32_t *cin = nullptr,
                     const uint32_t *ccin = nullptr, uint32_t *cout = nullptr,
                     uint32_t *ccout = nullptr, const LevelSide *side = nullptr,
                     const uint32_t *fsum = nullptr);

struct BottomUpArgs {
  const uint32_t *llist, *loff;  // the probe's long lists, per wave region,
  uint32_t nregions;             //   and the regions' bases (loff[nregions]: total)
  const uint32_t *rbit;       // residency, for the records-read counts
  const uint64_t *eptr;
  const uint32_t *deg;
  const uint32_t *front;      // current frontier
  uint32_t *vis, *next, *val;
  LevelCtr *rd_ctr;           // receives rd[2]
  uint32_t level_value;
  LevelCtr *nctr;             // next-frontier bookkeeping (see launch_bu_probe)
  LevelSide side;             // published after the kernel (no zero words here)
  const uint32_t *fsum = nullptr;  // frontier summary (launch_fsum), nullable
};
void launch_bottom_up(Graph &g, const BottomUpArgs &a, hipStream_t s);
// The BFS expansion of the frontier {v} (1 <= degree < 2^32) with the
// next-frontier sums (args as for A_BFS_NEXT): one launch, no queue; it
// also writes ctr's totals for {v} and val[v] = v_value.
void launch_expand_one(Graph &g, const ExpandArgs &args, uint32_t v, LevelCtr *ctr,
                       uint32_t v_value, hipStream_t s);
// One wave copies side's counters to the host-mapped readback (no zero
// words), after the launches before it on the stream.
void launch_publish(const LevelSide &side, hipStream_t s);
void build_parent_hints(Graph &g);
ExpandArgs expand_args(const Graph &g, const QueueSet &qs);
inline uint64_t bitmap_words(uint64_t n) { return n / 32 + 2; }
// Frontier summary: bit j of fsum is set when vertices [256 j, 256 j + 256)
// hold a frontier vertex (frontier words [8 j, 8 j + 8) not all zero).  A
// bottom-up scan given it (the short-list scan, the early-exit kernel) reads
// a frontier word only where the summary bit is set: for a small frontier
// most records are settled by the summary, which stays in L2.
constexpr uint32_t FSUM_SHIFT = 8;
inline uint64_t fsum_bits(uint64_t n) { return (bitmap_words(n) + 7) / 8; }
inline uint64_t fsum_words(uint64_t n) { return fsum_bits(n) / 32 + 3; }
void launch_fsum(Graph &g, const uint32_t *front, uint32_t *fsum, hipStream_t s);
void launch_expand(Graph &g, Algo a, const ExpandArgs &args, hipStream_t s);

// algos.hip
constexpr uint32_t PR_HUBS = 2048;       // hub table entries
constexpr uint32_t PR_HUB_SLOTS = 4096;  // LDS hash slots (power of two)
void ensure_pr_hubs(Graph &g);
void run_bfs(Graph &g, uint32_t src, uint32_t *out, thv_stats *st);
// Top-down BFS levels over a workspace the caller lays out (no allocation
// here); w.h_ctr must point at 2 host LevelCtr, w.ev may be null.
void bfs_topdown_in(Graph &g, Workspace &w, uint32_t src, thv_stats *st);
// noalloc.hip: BFS with its workspace in the library's static device array
void run_bfs_noalloc(Graph &g, uint32_t src, uint32_t *out, thv_stats *st);
void run_cc(Graph &g, uint32_t *out, thv_stats *st);
// cc.hip: components of a symmetric graph into ws.val; false when a link
// hit its step guard (the caller then uses label propagation)
bool cc_union_find(Graph &g, thv_stats *st);
void run_sssp(Graph &g, uint32_t src, uint32_t *out, thv_stats *st);
void run_pr(Graph &g, int mode, float damping, float tol, int max_iter,
            float *out, thv_stats *st);
void launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);

// rmat.hip
Graph *rmat_generate(int scale, int edge_factor, double a, double b, double c,
                     uint64_t seed, bool symmetric, bool weights,
                     bool transpose);

// batch.hip
void run_batch(Graph &g, int algo, const uint32_t *srcs, uint32_t k,
               uint64_t *hash, uint64_t *reached, uint32_t *depth,
               uint32_t *full, thv_stats *st);

// autotune.cpp
void autotune(Graph &g, const char *cache_path, int *chosen);
void tune(Graph &g, const thv_tune_cfg &cfg, thv_tune_result *out);

// shared hash (identical to the oracle's): SplitMix64 output function
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
// Entries of the degree and parent-hint arrays: n rounded up to whole tiles,
// so the probe sweep loads every lane of a sub-tile unconditionally (the
// padding holds degree 0 and no hint)
inline uint64_t vertex_pad(uint64_t n) { return div_up(n, TILE) * TILE; }

}  // namespace thv
