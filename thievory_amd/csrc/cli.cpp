// cli.cpp -- drop-in replacement for the reference's `main` binary.
//
// Same flag pairs and defaults as src/main.cu:16-42, 91-110:
//   --input F  --algo bfs|cc|sssp|pr  --edgeSize 4|8  --type push|pull
//   --source S (default 1)  --runs R (default 1)  --gpus G  --device D
// plus this engine's store flags:
//   --budget F        fraction of the edge bytes kept in HBM (default: 90%
//                     of free device memory, like graph.cuh:306-341)
//   --partition E     edges per partition (default 65536)
//   --tol T --maxIter I   PageRank stop rule (default 1e-7, 100)
//   --autotune FILE   pick the launch shape (cached in FILE)
//   --symmetric 0|1   the file holds every edge both ways (undirected
//                     input): BFS may run bottom-up levels, CC runs by
//                     union-find; results are the same (default 0)
// Output: the reference's stdout lines (graph.cuh:712-713, bfs/bfs.cu:
// 527-538) and results/values.bin (n x u32, or n x f64 for PageRank,
// graph.cuh:817-852; skipped silently when results/ does not exist).
//
// Host functions BFS32 / CC32 / SSSP32 / PR32 / PR32_PUSH mirror the
// reference's entry points (bfs/bfs.cuh:8, cc/cc.cuh:8, sssp/sssp.cuh:8,
// pr/pr.cuh:8, 13) without the NUMA map; they use only the C ABI (thv.h).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "thv.h"

namespace {

struct Options {
  std::string input, algo = "bfs", type = "push", tune;
  unsigned edge_size = 4, source = 1, runs = 1, gpus = 0;
  double budget = -1.0;
  unsigned partition = 1u << 16;
  float tol = 1e-7f;
  int max_iter = 100;
  bool symmetric = false;
};

void usage(const char *prog) {
  std::cout << "Usage: " << prog << std::endl
            << "Options:" << std::endl
            << "  --input : Specify the path to graph input " << std::endl
            << "  --algo : Specify the algorithm {'bfs', 'cc', 'pr', 'sssp'} (default = 'bfs')"
            << std::endl
            << "  --edgeSize : Specify the edge size {4, 8} (default = 4)" << std::endl
            << "  --type : Specify PageRank type {'push', 'pull'} (default = 'push')" << std::endl
            << "  --source : Specify the source vertex {0, 1, 2, ...} (default = 1)" << std::endl
            << "  --runs : Specify the number of runs {0, 1, 2, ...} (default = 1) " << std::endl
            << "  --gpus : Specify the number of NEIGHBOR GPUs {0} (default = 0)" << std::endl
            << "  --budget : Fraction of edge bytes kept in HBM (default = 90% of free)"
            << std::endl
            << "  --partition : Edges per partition (default = 65536)" << std::endl
            << "  --symmetric : The graph holds every edge both ways {0, 1} (default = 0)"
            << std::endl;
  exit(0);
}

[[noreturn]] void die(int rc, const char *what) {
  std::cerr << "Error: " << what << ": " << thv_last_error() << std::endl;
  exit(1);
}

thv_graph *open_graph(const Options &o, int fmt, int edge_bytes) {
  thv_graph *g = nullptr;
  int rc = thv_graph_open(o.input.c_str(), fmt, edge_bytes, &g);
  if (rc) die(rc, "cannot load the graph");
  uint64_t n = 0, m = 0;
  thv_graph_info(g, &n, &m, nullptr, nullptr);
  std::cout << "Num Vertices: " << n << std::endl << "Num Edges: " << m << std::endl;
  if (o.symmetric && fmt != THV_FMT_CSC &&
      (rc = thv_graph_set_option(g, THV_OPT_SYMMETRIC, 1.0)))
    die(rc, "cannot mark the graph symmetric");
  if (o.budget >= 0.0) {
    const uint64_t b = (uint64_t)(o.budget * (double)m * edge_bytes);
    if ((rc = thv_store_config(g, b, o.partition))) die(rc, "store configuration failed");
  } else if (o.partition != (1u << 16)) {
    uint64_t fr = 0;
    thv_device_memory(&fr, nullptr);
    if ((rc = thv_store_config(g, (uint64_t)(fr * 0.9), o.partition)))
      die(rc, "store configuration failed");
  }
  uint32_t P = 0, Pd = 0;
  uint64_t de = 0;
  thv_store_info(g, &P, &Pd, &de);
  std::cout << "Maximum number of edges per partition: " << o.partition << std::endl
            << "Number of partitions: " << P << std::endl
            << "Number of static edges: " << de << std::endl;
  if (!o.tune.empty()) {
    int ch[2];
    if ((rc = thv_autotune(g, o.tune.c_str(), ch))) die(rc, "autotune failed");
    std::cout << "Autotuned launch shape: " << ch[0] << " blocks/CU, " << ch[1]
              << " edges/thread" << std::endl;
  }
  return g;
}

template <class T>
void dump(const std::vector<T> &v) {
  std::ofstream f("results/values.bin", std::ios::binary);
  if (!f.is_open()) return;  // like CSR::DumpValues
  f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

template <class Run>
void timed_runs(const Options &o, uint64_t n, Run run) {
  double total = 0, host = 0;
  std::cout << "Starting Traversals" << std::endl;
  for (unsigned r = 0; r < o.runs; r++) {
    thv_stats st;
    run(st);
    std::cout << "Execution time: " << st.ms << " ms" << std::endl;
    total += st.ms;
    host += (double)st.host_bytes;
    if (r + 1 == o.runs)
      std::cout << "Edges traversed: " << st.edges_traversed << " ("
                << (st.ms > 0 ? st.edges_traversed / st.ms / 1e6 : 0.0) << " GTEPS), "
                << st.iterations << " iterations" << std::endl;
  }
  (void)n;
  if (o.runs) {
    std::cout << "Total amount of data read from host memory: " << host / 1048576.0 << " MB"
              << std::endl;
    std::cout << "Average execution time: " << total / o.runs << " ms" << std::endl;
  }
}

}  // namespace

// ---- reference-named entry points ------------------------------------------

void BFS32(const Options &o) {
  thv_graph *g = open_graph(o, THV_FMT_CSR, o.edge_size == 8 ? 8 : 4);
  uint64_t n = 0;
  thv_graph_info(g, &n, nullptr, nullptr, nullptr);
  std::vector<uint32_t> out(n);
  timed_runs(o, n, [&](thv_stats &st) {
    int rc = thv_bfs(g, o.source, out.data(), &st);
    if (rc) die(rc, "bfs failed");
  });
  dump(out);
  thv_graph_close(g);
}

void CC32(const Options &o) {
  thv_graph *g = open_graph(o, THV_FMT_CSR, o.edge_size == 8 ? 8 : 4);
  uint64_t n = 0;
  thv_graph_info(g, &n, nullptr, nullptr, nullptr);
  std::vector<uint32_t> out(n);
  timed_runs(o, n, [&](thv_stats &st) {
    int rc = thv_cc(g, out.data(), &st);
    if (rc) die(rc, "cc failed");
  });
  dump(out);
  thv_graph_close(g);
}

void SSSP32(const Options &o) {
  // weighted records: {dst, w} from a .wcsr, or the trailing weight array of
  // a .csr file (converter.cpp:771-772)
  const bool wcsr = o.input.size() >= 5 && o.input.compare(o.input.size() - 5, 5, ".wcsr") == 0;
  thv_graph *g = open_graph(o, wcsr ? THV_FMT_WCSR : THV_FMT_CSR, 8);
  uint64_t n = 0;
  thv_graph_info(g, &n, nullptr, nullptr, nullptr);
  std::vector<uint32_t> out(n);
  timed_runs(o, n, [&](thv_stats &st) {
    int rc = thv_sssp(g, o.source, out.data(), &st);
    if (rc) die(rc, "sssp failed");
  });
  dump(out);
  thv_graph_close(g);
}

static void PR_common(const Options &o, int mode) {
  thv_graph *g = open_graph(o, mode == THV_PR_PULL ? THV_FMT_CSC : THV_FMT_CSR, 4);
  uint64_t n = 0;
  thv_graph_info(g, &n, nullptr, nullptr, nullptr);
  std::vector<float> out(n);
  timed_runs(o, n, [&](thv_stats &st) {
    int rc = thv_pr(g, mode, 0.85f, o.tol, o.max_iter, out.data(), &st);
    if (rc) die(rc, "pagerank failed");
  });
  std::vector<double> d(out.begin(), out.end());  // values.bin holds f64
  dump(d);
  thv_graph_close(g);
}

void PR32(const Options &o) { PR_common(o, THV_PR_PULL); }
void PR32_PUSH(const Options &o) { PR_common(o, THV_PR_PUSH); }

int main(int argc, char **argv) {
  Options o;
  bool has_input = false;
  for (int i = 1; i < argc - 1; i += 2) {
    const char *k = argv[i], *v = argv[i + 1];
    if (!strcmp(k, "--input")) { o.input = v; has_input = true; }
    else if (!strcmp(k, "--algo")) o.algo = v;
    else if (!strcmp(k, "--edgeSize")) o.edge_size = (unsigned)atoi(v);
    else if (!strcmp(k, "--type")) o.type = v;
    else if (!strcmp(k, "--source")) o.source = (unsigned)atoi(v);
    else if (!strcmp(k, "--runs")) o.runs = (unsigned)atoi(v);
    else if (!strcmp(k, "--gpus")) o.gpus = (unsigned)atoi(v);
    else if (!strcmp(k, "--device")) { /* parsed and unused, as in main.cu:108-109 */ }
    else if (!strcmp(k, "--budget")) o.budget = atof(v);
    else if (!strcmp(k, "--partition")) o.partition = (unsigned)atoi(v);
    else if (!strcmp(k, "--tol")) o.tol = (float)atof(v);
    else if (!strcmp(k, "--maxIter")) o.max_iter = atoi(v);
    else if (!strcmp(k, "--autotune")) o.tune = v;
    else if (!strcmp(k, "--symmetric")) o.symmetric = atoi(v) != 0;
  }
  if (!has_input) exit(0);  // main.cu:116
  char arch[64] = "";
  int cus = 0;
  if (thv_device_info(arch, sizeof arch, &cus, nullptr) == 0)
    std::cout << "Selected device 0 (" << arch << ", " << cus << " CUs)" << std::endl;
  if (o.gpus != 0)
    std::cout << "Neighbor-GPU stealing is not part of this engine; running on one GPU"
              << std::endl;
  std::cout << "Running " << o.algo << " with edge size of " << o.edge_size
            << "B and using 0 neighbor GPUs" << std::endl;
  if (o.edge_size != 4 && o.edge_size != 8) {
    std::cout << "Error: wrong --edgeSize" << std::endl;
    usage(argv[0]);
  }
  if (o.algo == "bfs") {
    std::cout << "Source vertex is: " << o.source << std::endl;
    BFS32(o);
  } else if (o.algo == "cc") {
    CC32(o);
  } else if (o.algo == "sssp") {
    std::cout << "Source vertex is: " << o.source << std::endl;
    SSSP32(o);
  } else if (o.algo == "pr") {
    // --edgeSize 8 runs the pull variant (main.cu:211-212)
    if (o.edge_size == 8 || o.type == "pull") {
      std::cout << "Running PageRank Pull Implementation" << std::endl;
      PR32(o);
    } else if (o.type == "push") {
      std::cout << "Running PageRank Push Implementation" << std::endl;
      PR32_PUSH(o);
    } else {
      std::cout << "Error: wrong --type" << std::endl;
      usage(argv[0]);
    }
  } else {
    usage(argv[0]);
  }
  return 0;
}
