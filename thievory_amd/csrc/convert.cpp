// convert.cpp -- thv_convert: text edge list -> .csr / .csc / .wcsr.
//
// Same input handling and byte layouts as the reference's converter
// (converter/converter.cpp:483-589 CSC, 676-808 CSR and WCSR):
//   * skip --skip lines, then one "src dst [w]" edge per line, parsed with
//     the same formatted stream extraction (so odd lines behave alike);
//   * duplicates and self-loops are kept; --undirected also stores dst->src;
//   * n = largest id + 1; adjacency sorted by (src, dst) for CSR/WCSR and by
//     (dst, src) for CSC with std::sort, as the reference does;
//   * .csr : u64 n, u64 m, u64 off[n], u32 dst[m], u32 w[m]
//     .csc : u64 n, u64 m, u64 inoff[n], u32 src[m], u32 outdeg[n]
//     .wcsr: u64 n, u64 m, u64 off[n], {u32 dst, u32 w}[m]
// Weights are read from a third column with --weighted 1 (the reference's
// compile-time WEIGHTED), else 20.  Output names replace the input's last
// three characters ("txt") by the format name; .wcsr files are named
// <base>.wcsr (the reference writes <base>..wcsr, converter.cpp:777).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Edge {
  uint64_t src, dst, w;
};

struct Opts {
  std::string input, type = "csr", out_base;
  unsigned skip = 4;
  bool weighted = false, undirected = false;
};

bool read_edges(const Opts &o, std::vector<Edge> &edges, uint64_t &n) {
  std::ifstream in(o.input);
  if (!in.is_open()) {
    std::cerr << "Error opening file for reading: " << o.input << std::endl;
    return false;
  }
  std::string line;
  for (unsigned i = 0; i < o.skip; i++) std::getline(in, line);
  uint64_t src = 0, dst = 0, w = 20, maxid = 0;
  std::stringstream ss;
  while (std::getline(in, line)) {
    ss.str("");
    ss.clear();
    ss << line;
    ss >> src;
    ss >> dst;
    if (o.weighted) ss >> w;
    edges.push_back({src, dst, w});
    if (o.undirected) edges.push_back({dst, src, w});
    maxid = std::max(maxid, std::max(src, dst));
  }
  n = maxid + 1;
  return true;
}

bool write_all(std::ofstream &f, const void *p, size_t bytes) {
  f.write(static_cast<const char *>(p), (std::streamsize)bytes);
  return (bool)f;
}

std::string out_name(const Opts &o, const char *ext) {
  if (!o.out_base.empty()) return o.out_base + "." + ext;
  const std::string &s = o.input;
  return (s.size() >= 3 ? s.substr(0, s.size() - 3) : s) + ext;
}

int write_csr(const Opts &o, std::vector<Edge> edges, uint64_t n, bool interleaved) {
  const uint64_t m = edges.size();
  std::vector<uint64_t> off(n, 0);
  {
    std::vector<uint64_t> deg(n, 0);
    for (const Edge &e : edges) deg[e.src]++;
    uint64_t run = 0;
    for (uint64_t v = 0; v < n; v++) {
      off[v] = run;
      run += deg[v];
    }
  }
  std::sort(edges.begin(), edges.end(), [](const Edge &a, const Edge &b) {
    return a.src == b.src ? a.dst < b.dst : a.src < b.src;
  });
  const std::string path = out_name(o, interleaved ? "wcsr" : "csr");
  std::ofstream f(path, std::ios::binary);
  if (!f.is_open()) {
    std::cerr << "Error opening file for writing: " << path << std::endl;
    return 1;
  }
  const uint64_t hdr[2] = {n, m};
  bool ok = write_all(f, hdr, 16) && write_all(f, off.data(), n * 8);
  if (interleaved) {
    std::vector<uint32_t> rec(2 * m);
    for (uint64_t i = 0; i < m; i++) {
      rec[2 * i] = (uint32_t)edges[i].dst;
      rec[2 * i + 1] = (uint32_t)edges[i].w;
    }
    ok = ok && write_all(f, rec.data(), rec.size() * 4);
  } else {
    std::vector<uint32_t> a(m);
    for (uint64_t i = 0; i < m; i++) a[i] = (uint32_t)edges[i].dst;
    ok = ok && write_all(f, a.data(), m * 4);
    for (uint64_t i = 0; i < m; i++) a[i] = (uint32_t)edges[i].w;
    ok = ok && write_all(f, a.data(), m * 4);
  }
  if (!ok) {
    std::cerr << "Error writing " << path << std::endl;
    return 1;
  }
  std::cout << "Wrote " << path << " (" << n << " vertices, " << m << " edges)" << std::endl;
  return 0;
}

int write_csc(const Opts &o, std::vector<Edge> edges, uint64_t n) {
  const uint64_t m = edges.size();
  std::vector<uint32_t> indeg(n, 0), outdeg(n, 0);
  for (const Edge &e : edges) {
    indeg[e.dst]++;
    outdeg[e.src]++;
  }
  std::vector<uint64_t> off(n);
  uint64_t run = 0;
  for (uint64_t v = 0; v < n; v++) {
    off[v] = run;
    run += indeg[v];
  }
  std::sort(edges.begin(), edges.end(), [](const Edge &a, const Edge &b) {
    return a.dst == b.dst ? a.src < b.src : a.dst < b.dst;
  });
  std::vector<uint32_t> src(m);
  for (uint64_t i = 0; i < m; i++) src[i] = (uint32_t)edges[i].src;
  const std::string path = out_name(o, "csc");
  std::ofstream f(path, std::ios::binary);
  if (!f.is_open()) {
    std::cerr << "Error opening file for writing: " << path << std::endl;
    return 1;
  }
  const uint64_t hdr[2] = {n, m};
  if (!(write_all(f, hdr, 16) && write_all(f, off.data(), n * 8) &&
        write_all(f, src.data(), m * 4) && write_all(f, outdeg.data(), n * 4))) {
    std::cerr << "Error writing " << path << std::endl;
    return 1;
  }
  std::cout << "Wrote " << path << " (" << n << " vertices, " << m << " edges)" << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  Opts o;
  for (int i = 1; i < argc - 1; i += 2) {
    const char *k = argv[i], *v = argv[i + 1];
    if (!strcmp(k, "--input")) o.input = v;
    else if (!strcmp(k, "--skip")) o.skip = (unsigned)atoi(v);
    else if (!strcmp(k, "--type")) o.type = v;
    else if (!strcmp(k, "--weighted")) o.weighted = atoi(v) != 0;
    else if (!strcmp(k, "--undirected")) o.undirected = atoi(v) != 0;
    else if (!strcmp(k, "--output")) o.out_base = v;
  }
  if (o.input.empty()) {
    std::cout << "Usage: " << argv[0]
              << " --input FILE.txt [--skip K (4)] [--type csr|csc|wcsr|all] [--weighted 0|1]"
                 " [--undirected 0|1] [--output BASE]"
              << std::endl;
    return 0;
  }
  if (o.type != "csr" && o.type != "csc" && o.type != "wcsr" && o.type != "all") {
    std::cerr << "unknown --type " << o.type << std::endl;
    return 1;
  }
  std::vector<Edge> edges;
  uint64_t n = 0;
  if (!read_edges(o, edges, n)) return 1;
  if (n >= 0xFFFFFFFFull) {
    std::cerr << "vertex ids must be below 2^32-1" << std::endl;
    return 1;
  }
  int rc = 0;
  if (o.type == "csr" || o.type == "all") rc |= write_csr(o, edges, n, false);
  if (o.type == "wcsr" || o.type == "all") rc |= write_csr(o, edges, n, true);
  if (o.type == "csc" || o.type == "all") rc |= write_csc(o, edges, n);
  return rc;
}
