// wide.hpp -- graphs with 64-bit vertex ids (wide.hip).
#pragma once
#include <vector>

#include "internal.hpp"

namespace thv {

struct Graph64 {
  uint64_t n = 0, k = 0, m = 0;  // vertices, listed vertices (with edges), edges
  int rec = 8;                   // 8: u64 dst; 16: {u64 dst, u32 w, pad}
  std::vector<uint64_t> h_ids, h_loff;
  HostPinned h_edges;            // every record, mapped
  DevMem d_ids, d_loff, d_eptr;  // sorted listed ids, list offsets, list addresses
  DevMem slab;                   // HBM-resident runs of lists
  uint32_t P = 0, Pdev = 0;
  uint64_t dev_edges = 0, budget = UINT64_MAX;
  DevMem val, q[2], ctr, inq;    // values, queues, counters, queued marks
  hipStream_t stream = nullptr;
  int device = 0, cus = 256;
  std::mutex mu;
  ~Graph64();
};

// ids: k strictly ascending ids < n; loff: k + 1 offsets from 0, each list
// non-empty; dst: loff[k] destinations < n; w: nullable weights.
Graph64 *graph64_from_lists(uint64_t n, uint64_t k, const uint64_t *ids, const uint64_t *loff,
                            const uint64_t *dst, const uint32_t *w);
void graph64_store(Graph64 &g, uint64_t budget, uint32_t part_edges);
// algo THV_ALGO_BFS / CC / SSSP; out (nullable) n values; digest3
// (nullable): sum of mix64(mix64(v) + value) over finite values, their
// count and the largest value
void run_wide(Graph64 &g, int algo, uint64_t src, uint64_t *out, thv_stats *st,
              uint64_t *digest3);
void wide_gather(Graph64 &g, const uint64_t *at, uint64_t k, uint64_t *out);

}  // namespace thv
