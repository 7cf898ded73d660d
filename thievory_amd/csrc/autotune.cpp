// autotune.cpp -- picks the expansion launch shape for a graph and store.
//
// Candidates: blocks per CU in {2, 4, 8} x edges per thread in {4, 8}
// (chunk = 256 * ept edges, so also the unit hubs are split into).  Each is
// timed on BFS from three fixed sources (median); the fastest wins.  Results
// are cached on the handle and, keyed by (device, n, m, record bytes,
// budget, partition size), in a text cache file so later runs skip timing.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "internal.hpp"

namespace thv {

static std::string tune_key(const Graph &g) {
  hipDeviceProp_t p;
  HIPC(hipGetDeviceProperties(&p, g.device));
  std::ostringstream k;
  k << p.gcnArchName << "/" << g.cus << " n=" << g.n << " m=" << g.m << " rec=" << g.rec
    << " budget=" << g.store.budget << " part=" << g.store.part_edges;
  return k.str();
}

static bool cache_lookup(const char *path, const std::string &key, Shape &out) {
  if (!path) return false;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    const size_t bar = line.rfind('|');
    if (bar == std::string::npos || line.substr(0, bar) != key) continue;
    int b = 0, e = 0;
    if (sscanf(line.c_str() + bar + 1, "%d %d", &b, &e) == 2 && b > 0 && (e == 4 || e == 8)) {
      out.blocks_per_cu = b;
      out.ept = e;
      return true;
    }
  }
  return false;
}

void autotune(Graph &g, const char *cache_path, int *chosen) {
  const std::string key = tune_key(g);
  Shape best;
  if (!cache_lookup(cache_path, key, best)) {
    const Shape cands[] = {{2, 8}, {4, 8}, {8, 8}, {4, 4}, {8, 4}};
    double best_ms = 1e300;
    const uint32_t srcs[3] = {1u % (uint32_t)std::max<uint64_t>(g.n, 1),
                              (uint32_t)(g.n / 3), (uint32_t)(g.n / 2)};
    for (const Shape &s : cands) {
      g.shape = s;
      workspace_ensure(g);
      std::vector<double> t;
      for (uint32_t src : srcs) {
        if (src >= g.n) continue;
        thv_stats st;
        run_bfs(g, src, nullptr, &st);
        t.push_back(st.ms);
      }
      if (t.empty()) break;
      std::sort(t.begin(), t.end());
      const double med = t[t.size() / 2];
      if (med < best_ms) {
        best_ms = med;
        best = s;
      }
    }
    if (cache_path) {
      std::ofstream f(cache_path, std::ios::app);
      f << key << "|" << best.blocks_per_cu << " " << best.ept << "\n";
    }
  }
  g.shape = best;
  workspace_ensure(g);
  g.tuned = true;
  if (chosen) {
    chosen[0] = best.blocks_per_cu;
    chosen[1] = best.ept;
  }
}

}  // namespace thv
