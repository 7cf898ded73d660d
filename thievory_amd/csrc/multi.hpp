// multi.hpp -- multi-source traversal engine: up to 64 BFS or SSSP queries
// share every level of one traversal.
//
//   BFS : per-vertex 64-bit masks -- seen[v] (lanes that reached v), F[v]
//         (lanes with v in their frontier), N[v] (next frontier).  An edge
//         v -> w hands w the lanes F[v] & ~seen[w] (atomicOr); each newly
//         reached (lane, vertex) pair adds mix64(w << 32 | value) to the
//         lane's digest, so no per-lane value array is needed.  On a
//         symmetric graph dense levels run bottom-up: each vertex ORs its
//         neighbours' F masks until it has every lane it misses.
//   SSSP: lane-packed distances dist[v * L + q]; one wave relaxes one edge
//         for all its active lanes at once (lane q = query q), and the ballot
//         of improved lanes becomes w's next-frontier mask.
// Lanes are independent, so a query can join a free lane between any two
// levels (continuous batching): its level values are offset by the level it
// joined at.  A finished BFS lane is recycled by clearing its bit from
// seen[]; a finished SSSP lane is digested and reset to n+1 in one pass.
#pragma once
#include <vector>

#include "internal.hpp"

namespace thv {

struct LaneResult {
  uint64_t tag;       // caller's id for the query
  uint32_t src;
  uint32_t depth;
  uint64_t reached;
  uint64_t hash;
};

struct MultiEngine {
  static constexpr uint32_t MAXL = 64;

  Graph &g;
  const int algo;        // THV_ALGO_BFS or THV_ALGO_SSSP
  const uint32_t L;      // lanes (<= 64)
  const bool keep_full;  // keep per-lane value arrays (lane-major, L x n)
  hipStream_t stream = nullptr;
  QueueSet qs;
  DevMem fb[2], F[2];    // frontier bitmap and lane masks, current / next
  DevMem tbits;          // BFS: frontier vertices of the lanes a split level runs top-down
  DevMem seen, dist, full;
  DevMem acc;            // u64: hash[64] cnt[64] levelmask res_hash[64] res_cnt[64] res_depth[64]
  unsigned long long *h_acc = nullptr;
  LevelCtr *h_ctr = nullptr;
  int cur = 0;
  uint32_t level = 0;
  thv_stats stats;

  // host lane state
  uint64_t active = 0;         // lanes traversing
  uint64_t pend_finish = 0;    // SSSP lanes done, digest pending
  uint64_t pend_clear = 0;     // BFS lanes done, seen bits to clear before an admission
  uint64_t finalizing = 0;     // lanes whose digest is in flight this level
  uint32_t voff[MAXL] = {};
  uint32_t last_val[MAXL] = {};
  uint32_t src_of[MAXL] = {};
  uint64_t tag_of[MAXL] = {};
  uint32_t fin_src[MAXL] = {};      // query being digested in a lane
  uint64_t fin_tag[MAXL] = {};
  uint64_t just_finished = 0;       // lanes that ended in the last level
  std::vector<std::pair<uint32_t, uint32_t>> admit_q;  // (lane, src)
  bool launched_expand = false;
  bool launched_finish = false;
  // direction of BFS levels on a symmetric graph: bottom-up when
  // BU_ALPHA x (frontier edges) exceeds mu_est, an upper bound on the edges
  // of the vertices some active lane has not reached (measured by the last
  // bottom-up level, reset to m by every admission), and the frontier holds
  // more than n/8 edges (a bottom-up level sweeps all n vertices)
  static constexpr uint64_t BU_ALPHA = 2;
  uint64_t mu_est = 0;
  bool level_bu = false;
  // Per-lane direction inside a bottom-up level: a lane whose frontier
  // holds fewer than n / LANE_TD_DIV vertices runs top-down (its vertices
  // are few, and in a bottom-up scan the lanes that cannot reach a vertex at
  // this level keep its scan going to the end of the list); the rest run
  // bottom-up with only their own lanes missing.  Measured on a resident
  // R-MAT 25 batch of 64: n/1024 11.1 ms, n/8192 9.1, n/16384 9.1, n/131072
  // 9.9, no split 11.8.
  static constexpr uint64_t LANE_TD_DIV = 8192;
  uint64_t fsz[MAXL] = {};       // vertices in each lane's frontier
  uint64_t cnt_prev[MAXL] = {};  // each lane's reached count at the last level
  uint64_t level_td_lanes = 0;   // lanes the current bottom-up level ran top-down

  MultiEngine(Graph &g, int algo, uint32_t lanes, bool keep_full);
  ~MultiEngine();
  MultiEngine(const MultiEngine &) = delete;
  MultiEngine &operator=(const MultiEngine &) = delete;

  static uint64_t bytes_needed(const Graph &g, int algo, uint32_t lanes, bool keep_full);
  uint32_t chunk() const;                  // edges per expansion chunk
  uint64_t free_mask() const;
  uint32_t free_lanes() const { return (uint32_t)__builtin_popcountll(free_mask()); }
  int admit(uint32_t src, uint64_t tag);   // -> lane, or -1 when none is free
  bool busy() const { return active || pend_finish || !admit_q.empty(); }
  void launch_level();                     // enqueue one level on `stream`
  // expansion of the compacted frontier, pushing the given lanes
  void launch_top_down(LevelCtr *c, uint64_t lanes = ~0ull);
  void launch_bottom_up(uint64_t lanes);   // sweep + long lists (BFS, symmetric)
  void finish_level(std::vector<LaneResult> &done);  // sync, harvest finished lanes
  void copy_full(uint32_t lane, uint32_t *out);      // value array of a lane (keep_full)
};

}  // namespace thv
