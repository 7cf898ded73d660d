// abi.cpp -- the extern "C" boundary of libthv_hip.so (include/thv.h).
// Every entry point converts exceptions to THV_E* codes and records the
// message for thv_last_error() in thread-local storage.
#include <new>

#include "internal.hpp"
#include "wide.hpp"

namespace thv {
void set_last_error(const std::string &m);
const char *last_error();
Graph *graph_open(const char *path, int fmt, int edge_bytes);
void link_probe(uint64_t bytes, double out[4]);
Graph *graph_from_arrays(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *dst,
                         const uint32_t *w, const uint32_t *outdeg, int edge_bytes);
void graph_export(const Graph &g, uint64_t *off, uint32_t *dst, uint32_t *w, uint32_t *outdeg);
}  // namespace thv

using namespace thv;

template <class F>
static int guarded(F f) {
  try {
    f();
    return THV_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return THV_ENOMEM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return THV_EHIP;
  } catch (...) {
    set_last_error("unknown error");
    return THV_EHIP;
  }
}

#define NEED(x)                                               \
  do {                                                        \
    if (!(x)) fail(THV_EINVAL, "invalid argument: " #x);      \
  } while (0)

extern "C" {

int thv_graph_open(const char *path, int fmt, int edge_bytes, thv_graph **out) {
  return guarded([&] {
    NEED(out);
    *out = reinterpret_cast<thv_graph *>(graph_open(path, fmt, edge_bytes));
  });
}

int thv_graph_from_arrays(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *dst,
                          const uint32_t *w, const uint32_t *outdeg, int edge_bytes,
                          thv_graph **out) {
  return guarded([&] {
    NEED(out);
    *out = reinterpret_cast<thv_graph *>(
        graph_from_arrays(n, m, off, dst, w, outdeg, edge_bytes));
  });
}

int thv_graph_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                   int symmetric, int weights, int transpose, thv_graph **out) {
  return guarded([&] {
    NEED(out);
    *out = reinterpret_cast<thv_graph *>(rmat_generate(scale, edge_factor, a, b, c, seed,
                                                       symmetric != 0, weights != 0,
                                                       transpose != 0));
  });
}

int thv_graph_set_option(thv_graph *gh, int option, double value) {
  return guarded([&] {
    NEED(gh);
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    // options that decide the parent hints may rebuild the store
    auto refresh = [&] {
      if (g->servers.load() > 0 && store_stale(*g))
        fail(THV_EBUSY, "stop the servers first: this change rebuilds the store");
    };
    switch (option) {
      case THV_OPT_SYMMETRIC: {
        if (g->csc && value != 0.0) fail(THV_EINVAL, "a CSC graph holds in-edges only");
        const bool old = g->sym;
        g->sym = value != 0.0;
        try { refresh(); } catch (...) { g->sym = old; throw; }
        store_refresh(*g);
        break;
      }
      case THV_OPT_DIRECTION: {
        const bool old = g->dir_opt;
        g->dir_opt = value != 0.0;
        try { refresh(); } catch (...) { g->dir_opt = old; throw; }
        store_refresh(*g);
        break;
      }
      case THV_OPT_PARENT_HINTS: {
        const bool old = g->parent_hints;
        g->parent_hints = value != 0.0;
        try { refresh(); } catch (...) { g->parent_hints = old; throw; }
        store_refresh(*g);
        break;
      }
      case THV_OPT_DO_ALPHA:
        NEED(value > 0.0);
        g->do_alpha = value;
        break;
      case THV_OPT_DO_BETA:
        NEED(value > 0.0);
        g->do_beta = value;
        break;
      case THV_OPT_DO_VERTEX_NS:
        NEED(value >= 0.0);
        g->do_vertex_ns = value;
        break;
      case THV_OPT_DO_LINK_GBPS:
        NEED(value > 0.0);
        g->do_link_gbps = value;
        break;
      case THV_OPT_CC_UNION: g->cc_union = value != 0.0; break;
      case THV_OPT_CC_UF_STEPS:
        NEED(value >= 1.0 && value <= 4294967295.0);
        g->uf_max_steps = (uint32_t)value;
        break;
      case THV_OPT_SSSP_DELTA:
        NEED(value >= 0.0 && value <= 4294967295.0);
        g->sssp_delta = (uint32_t)value;
        break;
      case THV_OPT_DO_PULL_DEGREE:
        NEED(value >= 0.0 && value <= 16.0);
        g->do_pull_degree = (uint32_t)value;
        break;
      case THV_OPT_KERNEL_TIMING: g->kernel_timing = value != 0.0; break;
      case THV_OPT_DO_SUMMARY:
        NEED(value >= 0.0);
        g->do_summary_ratio = value;
        break;
      case THV_OPT_STORE_LAYOUT: {
        NEED(value == 0.0 || value == 1.0 || value == 2.0);
        const int old = g->store_layout;
        g->store_layout = (int)value;
        try { refresh(); } catch (...) { g->store_layout = old; throw; }
        store_refresh(*g);
        break;
      }
      default: fail(THV_EINVAL, "unknown option");
    }
  });
}

int thv_graph_get_option(const thv_graph *gh, int option, double *value) {
  return guarded([&] {
    NEED(gh && value);
    const Graph *g = reinterpret_cast<const Graph *>(gh);
    switch (option) {
      case THV_OPT_SYMMETRIC: *value = g->sym ? 1.0 : 0.0; break;
      case THV_OPT_DIRECTION: *value = g->dir_opt ? 1.0 : 0.0; break;
      case THV_OPT_DO_ALPHA: *value = g->do_alpha; break;
      case THV_OPT_DO_BETA: *value = g->do_beta; break;
      case THV_OPT_DO_VERTEX_NS: *value = g->do_vertex_ns; break;
      case THV_OPT_DO_LINK_GBPS: *value = g->do_link_gbps; break;
      case THV_OPT_DO_PULL_DEGREE: *value = g->do_pull_degree; break;
      case THV_OPT_PARENT_HINTS: *value = g->parent_hints ? 1.0 : 0.0; break;
      case THV_OPT_HINT_BYTES: *value = (double)g->store.d_hint.bytes; break;
      case THV_OPT_SSSP_DELTA: *value = g->sssp_delta; break;
      case THV_OPT_CC_UNION: *value = g->cc_union ? 1.0 : 0.0; break;
      case THV_OPT_CC_UF_STEPS: *value = g->uf_max_steps; break;
      case THV_OPT_KERNEL_TIMING: *value = g->kernel_timing ? 1.0 : 0.0; break;
      case THV_OPT_STORE_LAYOUT: *value = g->store_layout; break;
      case THV_OPT_DO_SUMMARY: *value = g->do_summary_ratio; break;
      default: fail(THV_EINVAL, "unknown option");
    }
  });
}

int thv_graph_info(const thv_graph *gh, uint64_t *n, uint64_t *m, int *edge_bytes, int *is_csc) {
  return guarded([&] {
    NEED(gh);
    const Graph *g = reinterpret_cast<const Graph *>(gh);
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (edge_bytes) *edge_bytes = g->rec;
    if (is_csc) *is_csc = g->csc ? 1 : 0;
  });
}

int thv_graph_export(const thv_graph *gh, uint64_t *off, uint32_t *dst, uint32_t *w,
                     uint32_t *outdeg) {
  return guarded([&] {
    NEED(gh);
    graph_export(*reinterpret_cast<const Graph *>(gh), off, dst, w, outdeg);
  });
}

int thv_graph_close(thv_graph *gh) {
  if (!gh) return THV_OK;
  Graph *g = reinterpret_cast<Graph *>(gh);
  {
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->servers.load() > 0) {
      set_last_error("a server is running on this graph: stop it first");
      return THV_EBUSY;
    }
  }
  try {
    delete g;
  } catch (...) {
  }
  return THV_OK;
}

int thv_store_config(thv_graph *gh, uint64_t budget, uint32_t partition_edges) {
  return guarded([&] {
    NEED(gh);
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->servers.load()) fail(THV_EBUSY, "a server is running on this graph");
    store_configure(*g, budget, partition_edges);
  });
}

int thv_store_info(const thv_graph *gh, uint32_t *partitions, uint32_t *device_partitions,
                   uint64_t *device_edges) {
  return guarded([&] {
    NEED(gh);
    Graph *g = const_cast<Graph *>(reinterpret_cast<const Graph *>(gh));
    {
      std::lock_guard<std::mutex> lk(g->mu);
      store_ensure(*g);
    }
    if (partitions) *partitions = g->store.P;
    if (device_partitions) *device_partitions = g->store.Pdev;
    if (device_edges) *device_edges = g->store.dev_edges;
  });
}

int thv_bfs(thv_graph *gh, uint32_t src, uint32_t *out, thv_stats *st) {
  return guarded([&] {
    NEED(gh);
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    run_bfs(*g, src, out, st);
  });
}

int thv_bfs_noalloc(thv_graph *gh, uint32_t src, uint32_t *out, thv_stats *st) {
  return guarded([&] {
    NEED(gh);
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    run_bfs_noalloc(*g, src, out, st);
  });
}

int thv_cc(thv_graph *gh, uint32_t *out, thv_stats *st) {
  return guarded([&] {
    NEED(gh);
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    run_cc(*g, out, st);
  });
}

int thv_sssp(thv_graph *gh, uint32_t src, uint32_t *out, thv_stats *st) {
  return guarded([&] {
    NEED(gh);
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    run_sssp(*g, src, out, st);
  });
}

int thv_pr(thv_graph *gh, int mode, float damping, float tol, int max_iter, float *out,
           thv_stats *st) {
  return guarded([&] {
    NEED(gh);
    NEED(damping > 0.0f && damping < 1.0f);
    NEED(tol >= 0.0f);
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    run_pr(*g, mode, damping, tol, max_iter, out, st);
  });
}

int thv_batch(thv_graph *gh, int algo, const uint32_t *srcs, uint32_t k, uint64_t *out_hash,
              uint64_t *out_reached, uint32_t *out_depth, uint32_t *out_full, thv_stats *st) {
  return guarded([&] {
    NEED(gh);
    NEED(srcs || k == 0);
    NEED(k <= 64);
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    run_batch(*g, algo, srcs, k, out_hash, out_reached, out_depth, out_full, st);
  });
}

int thv_graph64_from_lists(uint64_t n, uint64_t k, const uint64_t *ids, const uint64_t *loff,
                           const uint64_t *dst, const uint32_t *w, thv_graph64 **out) {
  return guarded([&] {
    NEED(out);
    *out = reinterpret_cast<thv_graph64 *>(graph64_from_lists(n, k, ids, loff, dst, w));
  });
}

int thv_graph64_store_config(thv_graph64 *gh, uint64_t budget, uint32_t partition_edges) {
  return guarded([&] {
    NEED(gh);
    Graph64 *g = reinterpret_cast<Graph64 *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    graph64_store(*g, budget, partition_edges);
  });
}

int thv_graph64_info(const thv_graph64 *gh, uint64_t *n, uint64_t *m, uint32_t *partitions,
                     uint32_t *device_partitions, uint64_t *device_edges) {
  return guarded([&] {
    NEED(gh);
    const Graph64 *g = reinterpret_cast<const Graph64 *>(gh);
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (partitions) *partitions = g->P;
    if (device_partitions) *device_partitions = g->Pdev;
    if (device_edges) *device_edges = g->dev_edges;
  });
}

static int wide_call(thv_graph64 *gh, int algo, uint64_t src, uint64_t *out, thv_stats *st,
                     uint64_t *digest3) {
  return guarded([&] {
    NEED(gh);
    Graph64 *g = reinterpret_cast<Graph64 *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    run_wide(*g, algo, src, out, st, digest3);
  });
}

int thv_bfs64(thv_graph64 *g, uint64_t src, uint64_t *out, thv_stats *st, uint64_t *digest3) {
  return wide_call(g, THV_ALGO_BFS, src, out, st, digest3);
}

int thv_cc64(thv_graph64 *g, uint64_t *out, thv_stats *st, uint64_t *digest3) {
  return wide_call(g, THV_ALGO_CC, 0, out, st, digest3);
}

int thv_sssp64(thv_graph64 *g, uint64_t src, uint64_t *out, thv_stats *st, uint64_t *digest3) {
  return wide_call(g, THV_ALGO_SSSP, src, out, st, digest3);
}

int thv_graph64_values(thv_graph64 *gh, const uint64_t *vertices, uint64_t k, uint64_t *out) {
  return guarded([&] {
    NEED(gh);
    if (k && (!vertices || !out)) fail(THV_EINVAL, "null argument");
    Graph64 *g = reinterpret_cast<Graph64 *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    wide_gather(*g, vertices, k, out);
  });
}

void thv_graph64_close(thv_graph64 *gh) {
  if (!gh) return;
  try {
    delete reinterpret_cast<Graph64 *>(gh);
  } catch (...) {
  }
}

int thv_link_probe(uint64_t bytes, double *out4) {
  return guarded([&] {
    NEED(out4);
    link_probe(bytes, out4);
  });
}

int thv_tune(thv_graph *gh, const thv_tune_cfg *cfg, thv_tune_result *out) {
  return guarded([&] {
    NEED(gh);
    if (!cfg) fail(THV_EINVAL, "null tuning configuration");
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->servers.load()) fail(THV_EBUSY, "a server is running on this graph");
    tune(*g, *cfg, out);
  });
}

int thv_autotune(thv_graph *gh, const char *cache_path, int *chosen) {
  return guarded([&] {
    NEED(gh);
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->servers.load()) fail(THV_EBUSY, "a server is running on this graph");
    autotune(*g, cache_path, chosen);
  });
}

const char *thv_last_error(void) { return last_error(); }

const char *thv_version(void) { return "thievory-amd 0.1 (gfx950)"; }

int thv_device_info(char *name, size_t name_len, int *cus, uint64_t *hbm_bytes) {
  return guarded([&] {
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPC(hipGetDeviceProperties(&p, dev));
    if (name && name_len) {
      strncpy(name, p.gcnArchName, name_len - 1);
      name[name_len - 1] = 0;
    }
    if (cus) *cus = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
  });
}

int thv_device_memory(uint64_t *free_bytes, uint64_t *total_bytes) {
  return guarded([&] {
    size_t f = 0, t = 0;
    HIPC(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
  });
}

}  // extern "C"
