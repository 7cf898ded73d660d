// autotune.cpp -- picks launch shapes, partition size and algorithm knobs
// for a graph, its store and a workload, by timing them on the GPU.
//
// The reference fixes these: 64 MiB partitions (PARTITION_SIZE_MB,
// src/common.cuh:39) and k * numSMs blocks of 1024 threads per kernel
// (src/bfs/bfs.cu:21-27).  Here each is a knob on the graph handle:
//   expansion grid (blocks per CU) x edges per thread     Shape
//   bottom-up probe / early-exit grid (blocks per CU)      Shape::probe_bpc
//   multi-source bottom-up grid (blocks per CU)           Shape::batch_bpc
//   bottom-up full-scan degree                            do_pull_degree
//   SSSP bucket width                                     sssp_delta
//   partition size (a store rebuild, same budget)         Store::part_edges
// thv_tune searches them one at a time from the handle's settings
// (coordinate descent: every candidate value of a knob is timed with the
// others fixed, the best is kept, then the next knob), timing the workload
// itself (median of a few runs) and checking that every candidate returns
// the starting configuration's results.  Choices are cached per (device,
// graph, store budget, workload) in a text file.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "internal.hpp"

namespace thv {

namespace {

struct Knobs {
  int bpc, ept, pbpc, bbpc;
  uint32_t part, delta, pull;
  int layout;  // THV_OPT_STORE_LAYOUT
};

Knobs current(const Graph &g) {
  return {g.shape.blocks_per_cu, g.shape.ept, g.shape.probe_bpc, g.shape.batch_bpc,
          g.store.part_edges,    g.sssp_delta, g.do_pull_degree, g.store_layout};
}

void apply(Graph &g, const Knobs &k) {
  store_ensure(g);
  g.store_layout = k.layout;
  if (k.part != g.store.part_edges || store_stale(g))
    store_configure(g, g.store.budget, k.part);
  g.shape.blocks_per_cu = k.bpc;
  g.shape.ept = k.ept;
  g.shape.probe_bpc = k.pbpc;
  g.shape.batch_bpc = k.bbpc;
  g.sssp_delta = k.delta;
  g.do_pull_degree = k.pull;
  workspace_ensure(g);
}

// Graph identity for the cache: sizes, layout, options and 64 sampled
// offsets (enough to tell two generated graphs of one size apart).
std::string fingerprint(const Graph &g) {
  uint64_t h = mix64(g.n ^ (g.m << 1));
  for (int i = 0; i <= 64 && !g.h_off.empty(); i++) {
    const uint64_t v = g.h_off.size() > 1 ? (uint64_t)i * (g.h_off.size() - 1) / 64 : 0;
    h = mix64(h ^ g.h_off[v] ^ ((uint64_t)i << 56));
  }
  std::ostringstream k;
  k << "n=" << g.n << " m=" << g.m << " rec=" << g.rec << " csc=" << g.csc << " sym=" << g.sym
    << " dir=" << g.dir_opt << " hints=" << g.parent_hints << " off=" << std::hex << h;
  return k.str();
}

std::string device_key(const Graph &g) {
  hipDeviceProp_t p;
  HIPC(hipGetDeviceProperties(&p, g.device));
  std::ostringstream k;
  k << p.gcnArchName << "/" << g.cus;
  return k.str();
}

const char *workload_name(int w) {
  switch (w) {
    case THV_TUNE_BFS: return "bfs";
    case THV_TUNE_SSSP: return "sssp";
    case THV_TUNE_PR: return "pr";
    case THV_TUNE_BATCH_BFS: return "batch-bfs";
    case THV_TUNE_BATCH_SSSP: return "batch-sssp";
  }
  return "?";
}

// Digest of a value array: the batch digest's rule (sum over finite values
// of mix64(v << 32 | value)), so that it is order-independent.
uint64_t digest(const uint32_t *x, uint64_t n, uint32_t none) {
  uint64_t h = 0;
  for (uint64_t v = 0; v < n; v++)
    if (x[v] != none) h += mix64((v << 32) | x[v]);
  return h;
}

struct Workload {
  int kind;
  std::vector<uint32_t> srcs;
  int repeats, pr_iter;
};

// One measurement: the median time over the repeats, and the result the
// check compares (a digest, or the PageRank vector).
struct Sample {
  double ms = 0;
  uint64_t dig = 0;
  std::vector<float> pr;
};

Sample measure(Graph &g, const Workload &w, bool check) {
  Sample out;
  std::vector<double> t;
  std::vector<uint32_t> vals;
  const uint32_t none = (uint32_t)(g.n + 1);
  const int mode = g.csc ? THV_PR_PULL : THV_PR_PUSH;
  for (int r = 0; r < w.repeats + (check ? 1 : 0); r++) {
    const bool keep = check && r == w.repeats;  // the last run returns results
    thv_stats st;
    double ms = 0;
    switch (w.kind) {
      case THV_TUNE_BFS:
      case THV_TUNE_SSSP:
        for (uint32_t s : w.srcs) {
          if (keep) vals.assign(g.n, 0);
          if (w.kind == THV_TUNE_BFS) run_bfs(g, s, keep ? vals.data() : nullptr, &st);
          else run_sssp(g, s, keep ? vals.data() : nullptr, &st);
          ms += st.ms;
          if (keep) out.dig = mix64(out.dig ^ digest(vals.data(), g.n, none));
        }
        break;
      case THV_TUNE_PR:
        if (keep) out.pr.assign(g.n, 0.0f);
        run_pr(g, mode, 0.85f, 0.0f, w.pr_iter, keep ? out.pr.data() : nullptr, &st);
        ms = st.ms;
        break;
      default: {  // batches
        const uint32_t k = (uint32_t)w.srcs.size();
        std::vector<uint64_t> hs(k), rc(k);
        std::vector<uint32_t> dp(k);
        run_batch(g, w.kind == THV_TUNE_BATCH_BFS ? THV_ALGO_BFS : THV_ALGO_SSSP, w.srcs.data(),
                  k, hs.data(), rc.data(), dp.data(), nullptr, &st);
        ms = st.ms;
        if (keep)
          for (uint32_t i = 0; i < k; i++) out.dig = mix64(out.dig ^ hs[i] ^ (rc[i] << 20) ^ dp[i]);
      }
    }
    if (!keep) t.push_back(ms);
  }
  std::sort(t.begin(), t.end());
  out.ms = t.empty() ? 0 : t[t.size() / 2];
  return out;
}

// PageRank results agree when their normalised L1 distance is at most 1e-5
// (float atomics add in a different order under another launch shape).
bool same_result(const Sample &a, const Sample &b) {
  if (a.pr.empty()) return a.dig == b.dig;
  double sa = 0, sb = 0, d = 0;
  for (size_t i = 0; i < a.pr.size(); i++) {
    sa += a.pr[i];
    sb += b.pr[i];
  }
  if (sa <= 0 || sb <= 0) return sa == sb;
  for (size_t i = 0; i < a.pr.size(); i++) d += std::fabs(a.pr[i] / sa - b.pr[i] / sb);
  return d <= 1e-5;
}

std::string knob_text(const Knobs &k, double d_ms, double t_ms) {
  std::ostringstream o;
  o.precision(17);  // the times read back exactly
  o << k.bpc << " " << k.ept << " " << k.pbpc << " " << k.bbpc << " " << k.part << " " << k.delta
    << " " << k.pull << " " << k.layout << " " << d_ms << " " << t_ms;
  return o.str();
}

bool cache_read(const char *path, const std::string &key, Knobs &k, double &d_ms, double &t_ms) {
  if (!path) return false;
  std::ifstream f(path);
  std::string line;
  bool found = false;
  while (std::getline(f, line)) {  // the last entry of a key wins
    const size_t bar = line.rfind('|');
    if (bar == std::string::npos || line.substr(0, bar) != key) continue;
    std::istringstream in(line.substr(bar + 1));
    Knobs x;
    double a = 0, b = 0;
    if (in >> x.bpc >> x.ept >> x.pbpc >> x.bbpc >> x.part >> x.delta >> x.pull >> x.layout >> a >>
            b &&
        x.bpc > 0 && (x.ept == 4 || x.ept == 8) && x.pbpc > 0 && x.bbpc > 0 && x.part > 0 &&
        x.pull <= 16 && x.layout >= 0 && x.layout <= 2) {
      k = x;
      d_ms = a;
      t_ms = b;
      found = true;
    }
  }
  return found;
}

std::vector<uint32_t> default_sources(const Graph &g, int kind) {
  // seeded, among vertices with edges; batches take 64, single-source runs 3
  const uint32_t want = (kind == THV_TUNE_BATCH_BFS || kind == THV_TUNE_BATCH_SSSP) ? 64 : 3;
  std::vector<uint32_t> s;
  uint64_t x = 0x5EEDull;
  for (uint32_t tries = 0; s.size() < want && tries < 64 * want && g.n; tries++) {
    x = mix64(x);
    const uint32_t v = (uint32_t)(x % g.n);
    if (g.h_off[v + 1] > g.h_off[v]) s.push_back(v);
  }
  if (s.empty() && g.n) s.push_back(0);
  return s;
}

// shape_only: the expansion shape alone (the first tuner's search).
void tune_impl(Graph &g, const thv_tune_cfg &cfg, thv_tune_result *res, bool shape_only) {
  const int kind = cfg.workload;
  if (kind < THV_TUNE_BFS || kind > THV_TUNE_BATCH_SSSP) fail(THV_EINVAL, "unknown workload");
  if ((kind == THV_TUNE_SSSP || kind == THV_TUNE_BATCH_SSSP) && g.rec != 8)
    fail(THV_ENOTSUP, "SSSP needs a weighted graph (8-byte records)");
  if (cfg.nsources > 64) fail(THV_EINVAL, "at most 64 timing sources");
  if (g.n == 0) fail(THV_EINVAL, "empty graph");
  Workload w;
  w.kind = kind;
  w.repeats = cfg.repeats > 0 ? cfg.repeats : 3;
  w.pr_iter = cfg.pr_iterations > 0 ? cfg.pr_iterations : 10;
  if (cfg.nsources) {
    if (!cfg.sources) fail(THV_EINVAL, "null sources");
    w.srcs.assign(cfg.sources, cfg.sources + cfg.nsources);
    for (uint32_t s : w.srcs)
      if (s >= g.n) fail(THV_EINVAL, "timing source out of range");
  } else {
    w.srcs = default_sources(g, kind);
  }
  store_ensure(g);
  std::ostringstream key;
  key << device_key(g) << " " << fingerprint(g) << " budget=" << g.store.budget
      << " work=" << workload_name(kind) << " parts=" << (cfg.partitions ? 1 : 0)
      << (shape_only ? " shape-only" : "");
  Knobs best = current(g);
  double d_ms = 0, t_ms = 0;
  uint32_t timed = 0;
  bool cached = cache_read(cfg.cache_path, key.str(), best, d_ms, t_ms);
  if (cached) {
    apply(g, best);
  } else {
    const bool bottom_up = g.sym && g.dir_opt;
    apply(g, best);
    const Sample base = measure(g, w, true);
    d_ms = t_ms = base.ms;
    timed = 1;
    // one knob at a time; a candidate must beat the incumbent by 2%
    auto try_knob = [&](const std::vector<Knobs> &cands) {
      for (const Knobs &c : cands) {
        apply(g, c);
        const Sample s = measure(g, w, true);
        timed++;
        if (!same_result(base, s))
          fail(THV_ESTATE, "tuning candidate changed the results (" +
                               knob_text(c, 0, 0) + "): engine bug");
        if (s.ms < 0.98 * t_ms) {
          t_ms = s.ms;
          best = c;
        }
      }
      apply(g, best);
    };
    auto vary = [&](auto set) {
      std::vector<Knobs> v;
      set(v);
      try_knob(v);
    };
    // first the store layout (the other of densest / sparsest partitions
    // first), which changes what every later knob is timed on
    if (!shape_only && cfg.partitions)
      vary([&](std::vector<Knobs> &v) {
        const bool hints = store_wants_hints(g, g.store.budget);
        const int was = g.store_layout;
        g.store_layout = best.layout;
        const bool cur = store_sparse_first(g, hints);
        for (int l : {0, 1, 2}) {
          g.store_layout = l;
          if (l != best.layout && store_sparse_first(g, hints) != cur) {
            Knobs k = best;
            k.layout = l;
            v.push_back(k);
            break;  // the other order: one candidate
          }
        }
        g.store_layout = was;
      });
    if (!shape_only && (kind == THV_TUNE_SSSP || kind == THV_TUNE_BATCH_SSSP))
      vary([&](std::vector<Knobs> &v) {
        for (uint32_t d : {8u, 16u, 32u, 64u})
          if (d != best.delta) {
            Knobs k = best;
            k.delta = d;
            v.push_back(k);
          }
      });
    vary([&](std::vector<Knobs> &v) {
      const int shapes[][2] = {{2, 8}, {4, 8}, {8, 8}, {4, 4}, {8, 4}};
      for (auto &sh : shapes)
        if (sh[0] != best.bpc || sh[1] != best.ept) {
          Knobs k = best;
          k.bpc = sh[0];
          k.ept = sh[1];
          v.push_back(k);
        }
    });
    if (!shape_only && kind == THV_TUNE_BFS && bottom_up) {
      vary([&](std::vector<Knobs> &v) {
        for (int p : {4, 8, 16})
          if (p != best.pbpc) {
            Knobs k = best;
            k.pbpc = p;
            v.push_back(k);
          }
      });
      vary([&](std::vector<Knobs> &v) {
        for (uint32_t d : {8u, 16u})
          if (d != best.pull) {
            Knobs k = best;
            k.pull = d;
            v.push_back(k);
          }
      });
    }
    if (!shape_only && (kind == THV_TUNE_BATCH_BFS || kind == THV_TUNE_BATCH_SSSP))
      vary([&](std::vector<Knobs> &v) {
        for (int p : {4, 8, 16})
          if (p != best.bbpc) {
            Knobs k = best;
            k.bbpc = p;
            v.push_back(k);
          }
      });
    if (!shape_only && cfg.partitions)
      vary([&](std::vector<Knobs> &v) {
        for (uint32_t p : {1u << 14, 1u << 16, 1u << 18})
          if (p != best.part) {
            Knobs k = best;
            k.part = p;
            v.push_back(k);
          }
      });
    if (cfg.cache_path) {
      std::ofstream f(cfg.cache_path, std::ios::app);
      f << key.str() << "|" << knob_text(best, d_ms, t_ms) << "\n";
    }
  }
  g.tuned = true;
  if (res) {
    memset(res, 0, sizeof *res);
    res->blocks_per_cu = best.bpc;
    res->edges_per_thread = best.ept;
    res->probe_blocks_per_cu = best.pbpc;
    res->batch_blocks_per_cu = best.bbpc;
    res->partition_edges = best.part;
    res->sssp_delta = best.delta;
    res->pull_degree = best.pull;
    res->candidates = cached ? 0 : timed;
    res->default_ms = d_ms;
    res->tuned_ms = t_ms;
    res->from_cache = cached ? 1 : 0;
    res->store_layout = best.layout;
  }
}

}  // namespace

void tune(Graph &g, const thv_tune_cfg &cfg, thv_tune_result *res) {
  tune_impl(g, cfg, res, false);
}

// The expansion shape alone, timed on BFS from three fixed sources (the
// first tuner; thv_tune covers it and more).
void autotune(Graph &g, const char *cache_path, int *chosen) {
  thv_tune_cfg c;
  memset(&c, 0, sizeof c);
  c.workload = THV_TUNE_BFS;
  c.repeats = 1;
  c.cache_path = cache_path;
  const uint32_t n = (uint32_t)std::max<uint64_t>(g.n, 1);
  const uint32_t srcs[3] = {1u % n, (uint32_t)(g.n / 3), (uint32_t)(g.n / 2)};
  c.sources = srcs;
  c.nsources = g.n ? 3 : 0;
  thv_tune_result r;
  tune_impl(g, c, &r, true);
  if (chosen) {
    chosen[0] = r.blocks_per_cu;
    chosen[1] = r.edges_per_thread;
  }
}

}  // namespace thv
