/*
 * thv.h -- C ABI of libthv_hip.so, the MI355X (gfx950) out-of-memory graph
 * engine.  Plain C: opaque handles, plain pointers and sizes, no C++ or torch
 * types.  Every entry point returns 0 (THV_OK) or a negative THV_E* code and
 * never lets a C++ exception escape; thv_last_error() gives the message of the
 * calling thread's last failure.
 *
 * Ownership: the library owns all device and pinned host memory; callers own
 * every output buffer they pass in.  A thv_graph is not safe for concurrent
 * algorithm calls (use one handle per thread); thv_serve_* is internally
 * synchronised.
 *
 * Value conventions follow the reference (qqueke/Thievory):
 *   BFS  out[v] = 1 + hops(src, v), out[src] = 1, unreached = n + 1
 *   CC   out[v] = least vertex id that reaches v along out-edges
 *   SSSP out[v] = min(n + 1, 1 + dist_w(src, v)), unreached = n + 1
 *   PR   out[v] = unnormalised PageRank (1-a) + a * sum r[u]/outdeg(u)
 * Citations are path:line in the reference.
 */
#ifndef THV_H
#define THV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define THV_OK 0
#define THV_EINVAL (-1)  /* bad argument (null handle, src >= n, k > 64 ...) */
#define THV_EIO (-2)     /* file cannot be opened or read */
#define THV_ENOMEM (-3)  /* device or pinned host allocation failed */
#define THV_EHIP (-4)    /* HIP runtime error */
#define THV_EFORMAT (-5) /* file size does not match its header */
#define THV_EBUSY (-6)   /* serving: no slot / scratch memory free */
#define THV_ENOTSUP (-7) /* operation not supported for this graph */
#define THV_ESTATE (-8)  /* handle in the wrong state (stopped server ...) */

/* input file formats (byte layouts of converter/converter.cpp) */
#define THV_FMT_CSR 0  /* u64 n,m; u64 off[n]; u32 dst[m]; [u32 w[m]]  converter.cpp:750-774 */
#define THV_FMT_CSC 1  /* u64 n,m; u64 inoff[n]; u32 src[m]; u32 outdeg[n]  converter.cpp:572-584 */
#define THV_FMT_WCSR 2 /* u64 n,m; u64 off[n]; {u32 dst,u32 w}[m]  converter.cpp:785-802 */

#define THV_ALGO_BFS 0
#define THV_ALGO_CC 1
#define THV_ALGO_SSSP 2
#define THV_ALGO_PR 3

#define THV_PR_PUSH 0
#define THV_PR_PULL 1

typedef struct thv_graph thv_graph;
typedef struct thv_server thv_server;

/* Counters of one algorithm call.  Byte figures are algorithmic: what the
 * algorithm must touch, not what the hardware moved (moved_bytes is the
 * exception: bytes the engine actually copied over the host link by DMA). */
typedef struct {
  uint64_t edges_traversed; /* sum over iterations of active out-edges   */
  uint64_t hbm_bytes;       /* device-resident edge records + state bytes */
  uint64_t host_bytes;      /* host-resident edge records read            */
  uint64_t moved_bytes;     /* bytes DMA'd host -> device                  */
  uint64_t reached;         /* vertices with a finite value               */
  uint32_t iterations;      /* levels / rounds until the frontier emptied */
  uint32_t host_partitions_touched; /* host partitions read, summed      */
  uint32_t host_partitions_skipped; /* host partitions skipped, summed    */
  uint32_t expand_launches;         /* edge-expansion kernel launches      */
  double ms;                /* first launch -> empty frontier (wall)      */
  double expand_ms;         /* time inside edge-expansion kernels (events)*/
  uint64_t edges_examined;  /* edge records actually read (bottom-up BFS
                               levels stop early, so this can be smaller
                               than edges_traversed)                       */
  uint32_t bottom_up_levels;/* BFS levels run bottom-up                    */
  uint32_t pad;
  uint64_t topdown_host_bytes; /* host_bytes / hbm_bytes a top-down-only    */
  uint64_t topdown_hbm_bytes;  /* traversal of the same levels would need   */
  double bottom_up_ms;         /* time inside bottom-up kernels (events);
                                  expand_ms counts the expansion kernel only */
  uint64_t bottom_up_host_bytes; /* part of host_bytes read bottom-up       */
  uint64_t hint_hits;            /* bottom-up BFS: vertices placed by their
                                    parent hint, without reading their list */
  double probe_ms;               /* bottom-up BFS: time inside the probe
                                    kernel (events), summed over its launches */
  uint32_t probe_launches;
  uint32_t pad2;
  uint64_t probe_hbm_bytes;      /* algorithmic HBM bytes of those launches */
} thv_stats;

/* ---- graph lifetime --------------------------------------------------- */

/* Opens a .csr/.csc/.wcsr file (graph.cuh:641-777).  edge_bytes 4 keeps
 * destination-only records; 8 keeps {dst, w} records (weights from a .wcsr
 * file or the trailing weight array of a .csr file; weight 20 if absent,
 * the converter's default, converter.cpp:695).  Replaces
 * CSR<uint32>::ReadInputFile + InitData. */
int thv_graph_open(const char *path, int fmt, int edge_bytes, thv_graph **out);

/* Builds a graph from host CSR arrays: off has n+1 entries, dst m, w m or
 * NULL.  edge_bytes as above.  outdeg (n, nullable) marks the arrays as a
 * CSC whose sources have these out-degrees (PageRank pull input). */
int thv_graph_from_arrays(uint64_t n, uint64_t m, const uint64_t *off,
                          const uint32_t *dst, const uint32_t *w,
                          const uint32_t *outdeg, int edge_bytes,
                          thv_graph **out);

/* Generates a seeded R-MAT graph on the GPU: 2^scale vertices, edge_factor
 * * 2^scale directed edges before clean-up (symmetric graphs draw half as
 * many tuples and store each both ways), self-loops and duplicates removed,
 * adjacency sorted.  weights != 0 attaches symmetric weights in [1, 64]
 * (edge_bytes 8).  transpose != 0 builds the CSC of the directed graph
 * instead (with out-degrees).  Bit-identical to the CPU generator. */
int thv_graph_rmat(int scale, int edge_factor, double a, double b, double c,
                   uint64_t seed, int symmetric, int weights, int transpose,
                   thv_graph **out);

int thv_graph_info(const thv_graph *g, uint64_t *n, uint64_t *m,
                   int *edge_bytes, int *is_csc);

/* Per-graph options (no reference counterpart).
 *   THV_OPT_SYMMETRIC   1: every edge u->v has v->u (set by thv_graph_rmat
 *                       for symmetric graphs; declare it for such files)
 *   THV_OPT_DIRECTION   1 (default): BFS on a symmetric graph switches to
 *                       bottom-up levels when the frontier is large
 *                       (direction-optimising BFS, Beamer et al. SC'12);
 *                       0: top-down only, as in the reference
 *   THV_OPT_DO_ALPHA    all edges in HBM: top-down -> bottom-up when
 *                       frontier edges exceed unvisited edges / alpha
 *                       (default 14)
 *   THV_OPT_DO_BETA     all edges in HBM: bottom-up -> top-down when the
 *                       frontier holds fewer than n / beta vertices (24)
 *   THV_OPT_DO_VERTEX_NS  split store: a level runs bottom-up when its
 *                       estimated link time is less than (frontier host
 *                       edge bytes / link rate); each unvisited vertex with
 *                       a long host list costs this many ns (one small
 *                       zero-copy read); default 3 (measured on MI355X)
 *   THV_OPT_DO_LINK_GBPS  the link rate of that rule, default 55
 *   THV_OPT_DO_PULL_DEGREE  a bottom-up level scans the whole list of every
 *                       unvisited vertex of at most this degree (0..16,
 *                       default 16) inside its probe sweep; longer lists
 *                       stop at the first frontier neighbour.  0 sends
 *                       every vertex to the early-exit scan.  The split-store rule then prices a
 *                       bottom-up level as (bytes of those short lists on
 *                       the host / link rate) + (longer unvisited host
 *                       lists x DO_VERTEX_NS).
 *   THV_OPT_PARENT_HINTS  1 (default): a symmetric graph with
 *                       direction-optimising BFS keeps one parent hint per
 *                       vertex in HBM -- its neighbour of largest degree.
 *                       A bottom-up level first tests the hints; a vertex
 *                       whose hint is in the frontier joins the next level
 *                       without its list being read.  A resident store
 *                       holds them beside the lists (n x 4 bytes more); a
 *                       split store with 4-byte records takes them from the
 *                       budget.  0: no hints, the whole budget holds
 *                       partitions.  Changing it, SYMMETRIC or DIRECTION
 *                       rebuilds the store with the same budget.
 *   THV_OPT_HINT_BYTES  (read only) HBM bytes the hints use, 0 if none.
 *   THV_OPT_SSSP_DELTA  SSSP bucket width: vertices wait in distance buckets
 *                       of this width before their edges are relaxed
 *                       (delta-stepping; default 16).  0: frontier
 *                       Bellman-Ford, the reference's order.  Distances are
 *                       the same either way.
 *   THV_OPT_CC_UNION    1 (default): CC of a symmetric graph by union-find
 *                       with neighbour sampling (the labels are the
 *                       component minima, as label propagation gives);
 *                       0: label propagation, as in the reference.
 *   THV_OPT_CC_UF_STEPS step guard of a union-find link or find (default
 *                       2^20, >= 1).  A link that needs more steps raises
 *                       a flag and CC reruns by label propagation; the
 *                       labels are the same either way.
 *   THV_OPT_KERNEL_TIMING 1 (default): BFS brackets its expansion, probe
 *                       and bottom-up launches with HIP events and reports
 *                       their time (thv_stats expand_ms, probe_ms,
 *                       bottom_up_ms); 0: no events, those stay 0 (each
 *                       event is a few microseconds of queue time).
 *   THV_OPT_STORE_LAYOUT which partitions a split store keeps in HBM: 0
 *                       (default) densest first, or sparsest first with the
 *                       short host lists copied to HBM when the graph holds
 *                       parent hints; 1 densest first; 2 sparsest first with
 *                       the short-list side area.  Changing it rebuilds the
 *                       store with the same budget.
 *   THV_OPT_DO_SUMMARY  a bottom-up level whose frontier holds fewer than
 *                       n / value vertices (default 1024; 0 never), at
 *                       the traversal's first probe, builds a frontier summary
 *                       (one bit per 256 vertices) and its scans test it
 *                       before reading a frontier word; a value below 1
 *                       builds it at every bottom-up level.
 * BFS results are the same either way (levels are unique). */
#define THV_OPT_SYMMETRIC 1
#define THV_OPT_DIRECTION 2
#define THV_OPT_DO_ALPHA 3
#define THV_OPT_DO_BETA 4
#define THV_OPT_DO_VERTEX_NS 5
#define THV_OPT_DO_LINK_GBPS 6
#define THV_OPT_DO_PULL_DEGREE 7
#define THV_OPT_PARENT_HINTS 8
#define THV_OPT_HINT_BYTES 9
#define THV_OPT_SSSP_DELTA 10
#define THV_OPT_CC_UNION 11
#define THV_OPT_CC_UF_STEPS 12
#define THV_OPT_KERNEL_TIMING 13
#define THV_OPT_STORE_LAYOUT 14
#define THV_OPT_DO_SUMMARY 15
int thv_graph_set_option(thv_graph *g, int option, double value);
int thv_graph_get_option(const thv_graph *g, int option, double *value);

/* Copies the CSR arrays to host buffers (any may be NULL). */
int thv_graph_export(const thv_graph *g, uint64_t *off, uint32_t *dst,
                     uint32_t *w, uint32_t *outdeg);

/* Frees the graph.  THV_EBUSY (and the graph stays open) while a server
 * runs on it. */
int thv_graph_close(thv_graph *g);

/* ---- split edge store ------------------------------------------------- */

/* Splits the edge array into partitions of at most partition_edges edges
 * (contiguous vertex ranges; one oversize vertex forms its own partition,
 * graph.cuh:270-304) and keeps the highest-average-degree partitions in HBM
 * up to device_edge_budget_bytes; the rest stays in mapped pinned host
 * memory read zero-copy.  Replaces CSR::SetPartitionsConfig and
 * CSR::SetNumStaticEdges (graph.cuh:306-341).  The default store (before any
 * call) is fully device-resident with 2^16-edge partitions. */
int thv_store_config(thv_graph *g, uint64_t device_edge_budget_bytes,
                     uint32_t partition_edges);

/* Store facts: partitions, partitions in HBM, edges in HBM. */
int thv_store_info(const thv_graph *g, uint32_t *partitions,
                   uint32_t *device_partitions, uint64_t *device_edges);

/* ---- single-source / whole-graph algorithms --------------------------- */

/* BFS levels (BFS32, bfs/bfs.cu:8-542).  out: n x u32 host buffer. */
int thv_bfs(thv_graph *g, uint32_t src, uint32_t *out, thv_stats *st);

/* BFS with no allocation (no counterpart in the reference; for callers that
 * cannot allocate: code under stream capture, tight serving loops).  The
 * same levels as thv_bfs, by top-down levels only, with every per-call
 * buffer (queues, tile tables, bitmaps, values, counters) carved from one
 * static device array of THV_NOALLOC_SCRATCH_BYTES per GPU, which the GPU
 * holds from library load to process end.  The call allocates no device or
 * pinned memory.  The graph's store must be built already (thv_store_config
 * or an earlier algorithm call), else THV_ESTATE.  The workspace takes about
 * 30 bytes per vertex plus 4 per partition and 8 per 2048 edges; a graph
 * that needs more than the array is THV_EINVAL.  Every call on one GPU
 * shares its array: calls that overlap on one GPU race.  out: n x u32 host
 * buffer; st as for thv_bfs (expand_ms stays 0: no events are created). */
#define THV_NOALLOC_SCRATCH_BYTES (64ull << 20)
int thv_bfs_noalloc(thv_graph *g, uint32_t src, uint32_t *out, thv_stats *st);

/* Min-label propagation (CC32, cc/cc.cu:71-614). */
int thv_cc(thv_graph *g, uint32_t *out, thv_stats *st);

/* Frontier Bellman-Ford with atomic min (SSSP32, sssp/sssp.cu:10-570), with
 * the graph's weights (edge_bytes 8). */
int thv_sssp(thv_graph *g, uint32_t src, uint32_t *out, thv_stats *st);

/* PageRank (PR32_PUSH pr/pr.cu:1030-1549 on a CSR graph, PR32 pull
 * pr/pr.cu:11-519 on a CSC graph), fp32 ranks, stop when no vertex moves by
 * more than tol or after max_iter iterations. */
int thv_pr(thv_graph *g, int mode, float damping, float tol, int max_iter,
           float *out, thv_stats *st);

/* ---- 64-bit vertex ids ------------------------------------------------- */

/* Graphs of 2^32 - 1 vertices or more: the reference declares BFS64, CC64
 * and SSSP64 (bfs/bfs.cu:544, cc/cc.cu:616, sssp/sssp.cu:572) but leaves
 * them empty, and its file format holds 4-byte destinations, so these
 * graphs enter through arrays.  The edge lists are given per listed vertex
 * (a vertex with edges): k strictly ascending ids < n, k + 1 list offsets
 * from 0 (every list non-empty) and loff[k] destinations < n, with
 * optional u32 weights (needed by SSSP).  Values are n x u64 with the
 * conventions above (unreached / none: n + 1).  The store is split as for
 * thv_store_config (greedy runs of lists of at most partition_edges
 * records, densest in HBM up to the budget, the rest zero-copy); the
 * default is all in HBM with 2^16-record runs.  Traversal is top-down.
 * digest3 (nullable) receives the sum of mix64(mix64(v) + value) over the
 * vertices with a value other than n + 1 (mod 2^64; CC: over all), their
 * count and the largest value; out (nullable) the n values. */
typedef struct thv_graph64 thv_graph64;
int thv_graph64_from_lists(uint64_t n, uint64_t k, const uint64_t *ids,
                           const uint64_t *loff, const uint64_t *dst,
                           const uint32_t *w, thv_graph64 **out);
int thv_graph64_store_config(thv_graph64 *g, uint64_t device_edge_budget_bytes,
                             uint32_t partition_edges);
int thv_graph64_info(const thv_graph64 *g, uint64_t *n, uint64_t *m,
                     uint32_t *partitions, uint32_t *device_partitions,
                     uint64_t *device_edges);
int thv_bfs64(thv_graph64 *g, uint64_t src, uint64_t *out, thv_stats *st,
              uint64_t *digest3);
int thv_cc64(thv_graph64 *g, uint64_t *out, thv_stats *st, uint64_t *digest3);
int thv_sssp64(thv_graph64 *g, uint64_t src, uint64_t *out, thv_stats *st,
               uint64_t *digest3);
/* The last algorithm call's values at k given vertices. */
int thv_graph64_values(thv_graph64 *g, const uint64_t *vertices, uint64_t k,
                       uint64_t *out);
void thv_graph64_close(thv_graph64 *g);

/* ---- batched multi-source queries (no reference counterpart) ---------- */

/* Runs k <= 64 BFS (algo THV_ALGO_BFS) or SSSP (THV_ALGO_SSSP) queries in
 * one persistent launch sequence: per-vertex 64-bit source masks for BFS,
 * lane-packed distances for SSSP.  out_hash[k] receives the digest
 *   sum over v with out[v] != n+1 of mix64((v << 32) | out[v])  (mod 2^64)
 * where mix64 is the SplitMix64 output function, out_reached[k] the reached
 * counts, out_depth[k] the largest finite value.  out_full (nullable) gets
 * the k full value arrays, query-major (k x n). */
int thv_batch(thv_graph *g, int algo, const uint32_t *srcs, uint32_t k,
              uint64_t *out_hash, uint64_t *out_reached, uint32_t *out_depth,
              uint32_t *out_full, thv_stats *st);

/* Chooses the expansion launch shape (blocks per CU, edges per thread) for
 * this graph by timing candidates; the choice is cached on the handle and
 * written to cache_path (nullable).  chosen receives 2 ints. */
int thv_autotune(thv_graph *g, const char *cache_path, int *chosen);

/* Tuning (no reference counterpart; it replaces the reference's fixed
 * PARTITION_SIZE_MB, common.cuh:39, and k*numSMs x 1024 launch shapes,
 * bfs/bfs.cu:21-27).  thv_tune times one workload on this graph under
 * candidate configurations, one knob at a time (coordinate descent from
 * the graph's current settings), and keeps the fastest:
 *   expansion grid x edges per thread (2/4/8 blocks per CU x 4/8),
 *   bottom-up probe grid (4/8/16 blocks per CU; BFS on symmetric graphs),
 *   multi-source grid (4/8/16 blocks per CU; batches),
 *   bottom-up full-scan degree (8/16; BFS on symmetric graphs),
 *   SSSP bucket width (8/16/32/64),
 *   with partitions != 0, each rebuilding the store under the same HBM
 *   budget: the store layout (the other of densest / sparsest partitions
 *   first, THV_OPT_STORE_LAYOUT) and the partition size (2^14, 2^16, 2^18
 *   edges).
 * Every candidate's results are compared with the starting configuration's
 * (bit-exact digests; PageRank within 1e-5 normalised L1): a difference is
 * THV_ESTATE.  The choice stays on the handle and, keyed by device, graph
 * (size, record bytes, options, sampled offsets), store budget and
 * workload, is appended to cache_path (nullable); a later call with the same
 * key applies the cached choice without timing. */
#define THV_TUNE_BFS 0
#define THV_TUNE_SSSP 1
#define THV_TUNE_PR 2          /* push on a CSR graph, pull on a CSC graph */
#define THV_TUNE_BATCH_BFS 3
#define THV_TUNE_BATCH_SSSP 4
typedef struct {
  int workload;              /* THV_TUNE_*                                  */
  uint32_t nsources;         /* timing sources, <= 64 (0: seeded defaults)  */
  const uint32_t *sources;
  int partitions;            /* 1: also try partition sizes (store rebuilds) */
  int repeats;               /* timed runs per candidate (0: 3)             */
  int pr_iterations;         /* PageRank iterations per timed run (0: 10)   */
  const char *cache_path;    /* nullable                                    */
} thv_tune_cfg;
typedef struct {
  int blocks_per_cu, edges_per_thread;  /* expansion shape                  */
  int probe_blocks_per_cu, batch_blocks_per_cu;
  uint32_t partition_edges, sssp_delta, pull_degree;
  uint32_t candidates;       /* configurations timed (0: from the cache)     */
  double default_ms;         /* median workload time, starting settings     */
  double tuned_ms;           /*   and chosen settings                        */
  int from_cache;
  int store_layout;          /* THV_OPT_STORE_LAYOUT chosen (with partitions) */
} thv_tune_result;
int thv_tune(thv_graph *g, const thv_tune_cfg *cfg, thv_tune_result *out);

/* ---- serving (continuous batching) ------------------------------------ */

typedef struct {
  uint32_t max_lanes;       /* concurrent queries per algorithm, <= 64     */
  uint64_t scratch_bytes;   /* device memory admission budget (0 = auto)   */
  uint32_t keep_full;       /* 1: keep full value arrays for thv_serve_poll */
  uint32_t sssp_engines;    /* 64-lane SSSP engines, each on its own stream
                               (0 = 1), while the scratch budget holds them */
} thv_serve_cfg;

typedef struct {
  uint64_t ticket;
  int32_t algo;
  int32_t status;           /* 0 done, 1 pending, negative error          */
  uint32_t src;
  uint32_t depth;
  uint64_t reached;
  uint64_t hash;
  double submit_ms, start_ms, done_ms; /* on the server's clock           */
} thv_result;

/* Starts a server thread that keeps the graph resident and admits queries
 * into free lanes between levels of the running batch. */
int thv_serve_start(thv_graph *g, const thv_serve_cfg *cfg, thv_server **out);
/* Thread-safe; returns THV_EBUSY when the admission queue is full. */
int thv_serve_submit(thv_server *s, int algo, uint32_t src, uint64_t *ticket);
/* Thread-safe; status 1 while pending.  full (nullable, n entries) receives
 * the value array when the server keeps them.  A finished query (status 0
 * or an error) is returned once: the poll that returns it releases its
 * entry and value array, and a later poll of the ticket is THV_EINVAL.
 * With keep_full, at most max(64, 4 GiB / (4 n)) value arrays (finished and
 * unpolled, plus queries still running) exist at once; submit returns
 * THV_EBUSY beyond that. */
int thv_serve_poll(thv_server *s, uint64_t ticket, thv_result *res,
                   uint32_t *full);
/* Blocks until every submitted query has finished. */
int thv_serve_drain(thv_server *s);
void thv_serve_stop(thv_server *s);

/* ---- misc ---------------------------------------------------------------- */

const char *thv_last_error(void);
/* Library version string and the device it runs on. */
const char *thv_version(void);
int thv_device_info(char *name, size_t name_len, int *cus,
                    uint64_t *hbm_bytes);
/* Ceilings of the current device (no reference counterpart), over buffers
 * of `bytes` bytes, best of three timed runs: out[0] a streaming kernel's
 * zero-copy read of mapped pinned host memory (the host share of the edge
 * store), out[1] pinned host -> device DMA, out[2] the same kernel's read of
 * HBM, out[3] a device -> device copy (bytes read + written); all GB/s. */
int thv_link_probe(uint64_t bytes, double *out4);

/* Free and total device memory (hipMemGetInfo). */
int thv_device_memory(uint64_t *free_bytes, uint64_t *total_bytes);

#ifdef __cplusplus
}
#endif
#endif /* THV_H */
