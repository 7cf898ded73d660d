// io.cpp -- reading .csr / .csc / .wcsr files and host arrays into a Graph.
//
// Byte layouts are those written by the reference converter
// (converter/converter.cpp:572-584 CSC, 750-774 CSR, 785-802 WCSR) and read
// by CSR::ReadInputFile (src/graph.cuh:707-777): u64 n, u64 m, u64 off[n]
// (off[n] = m is implied), then the records.  Unlike the reference, the
// file size is checked against the header, offsets must be monotone and
// every endpoint must be < n; violations fail with THV_EFORMAT instead of
// reading out of bounds.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <thread>

#include "internal.hpp"

namespace thv {

namespace {

struct Fd {
  int fd = -1;
  ~Fd() {
    if (fd >= 0) close(fd);
  }
};

void pread_all(int fd, void *dst, uint64_t bytes, uint64_t off, const char *path) {
  char *p = static_cast<char *>(dst);
  while (bytes) {
    ssize_t r = pread(fd, p, std::min<uint64_t>(bytes, 1ull << 30), (off_t)off);
    if (r <= 0) fail(THV_EIO, std::string("short read from ") + path);
    p += r;
    off += (uint64_t)r;
    bytes -= (uint64_t)r;
  }
}

template <class F>
void parallel_for(uint64_t n, F f) {
  unsigned t = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < (1u << 20)) t = 1;
  std::vector<std::thread> th;
  for (unsigned i = 0; i < t; i++)
    th.emplace_back([=] { f(n * i / t, n * (i + 1) / t); });
  for (auto &x : th) x.join();
}

void check_offsets(const std::vector<uint64_t> &off, uint64_t n, uint64_t m) {
  if (n && off[0] != 0) fail(THV_EFORMAT, "first offset is not 0");
  for (uint64_t v = 0; v < n; v++)
    if (off[v] > off[v + 1]) fail(THV_EFORMAT, "offsets decrease at vertex " + std::to_string(v));
  if (n && off[n - 1] > m) fail(THV_EFORMAT, "offsets exceed the edge count");
}

// Every endpoint (at stride `stride` u32 words) must be < n.
void check_endpoints(const uint32_t *p, uint64_t m, int stride, uint64_t n) {
  std::atomic<uint64_t> bad{UINT64_MAX};
  parallel_for(m, [&](uint64_t a, uint64_t b) {
    for (uint64_t i = a; i < b; i++)
      if (p[i * stride] >= n) {
        uint64_t cur = bad.load();
        while (i < cur && !bad.compare_exchange_weak(cur, i)) {
        }
        return;
      }
  });
  if (bad.load() != UINT64_MAX)
    fail(THV_EFORMAT, "edge " + std::to_string(bad.load()) + " names a vertex >= n");
}

}  // namespace

Graph *graph_open(const char *path, int fmt, int edge_bytes) {
  if (!path) fail(THV_EINVAL, "null path");
  if (edge_bytes != 4 && edge_bytes != 8) fail(THV_EINVAL, "edge_bytes must be 4 or 8");
  if (fmt != THV_FMT_CSR && fmt != THV_FMT_CSC && fmt != THV_FMT_WCSR)
    fail(THV_EINVAL, "unknown file format");
  if (fmt == THV_FMT_CSC && edge_bytes != 4)
    fail(THV_EINVAL, "a CSC file has 4-byte records");
  Fd f;
  f.fd = open(path, O_RDONLY);
  if (f.fd < 0) fail(THV_EIO, std::string("cannot open ") + path);
  struct stat sb;
  if (fstat(f.fd, &sb) != 0) fail(THV_EIO, std::string("cannot stat ") + path);
  const uint64_t size = (uint64_t)sb.st_size;
  uint64_t hdr[2];
  if (size < 16) fail(THV_EFORMAT, std::string(path) + " is shorter than its header");
  pread_all(f.fd, hdr, 16, 0, path);
  const uint64_t n = hdr[0], m = hdr[1];
  if (n >= 0xFFFFFFFFull) fail(THV_ENOTSUP, "graphs need fewer than 2^32-1 vertices");
  if (n > (size / 8) || m > size) fail(THV_EFORMAT, "header does not match the file size");
  const uint64_t base = 16 + 8 * n;
  bool file_w = false;
  if (fmt == THV_FMT_CSR) {
    if (size == base + 8 * m) file_w = true;
    else if (size != base + 4 * m) fail(THV_EFORMAT, "file size does not match a .csr layout");
  } else if (fmt == THV_FMT_CSC) {
    if (size != base + 4 * m + 4 * n) fail(THV_EFORMAT, "file size does not match a .csc layout");
  } else {
    if (size != base + 8 * m) fail(THV_EFORMAT, "file size does not match a .wcsr layout");
    file_w = true;
  }

  std::unique_ptr<Graph> g(new Graph());
  graph_init_device(*g);
  g->n = n;
  g->m = m;
  g->rec = edge_bytes;
  g->csc = fmt == THV_FMT_CSC;
  g->h_off.resize(n + 1);
  if (n) pread_all(f.fd, g->h_off.data(), 8 * n, 16, path);
  g->h_off[n] = m;
  check_offsets(g->h_off, n, m);

  g->h_edges.alloc(m * (uint64_t)edge_bytes);
  uint32_t *rec = static_cast<uint32_t *>(g->h_edges.p);
  if (fmt == THV_FMT_WCSR && edge_bytes == 8) {
    pread_all(f.fd, rec, 8 * m, base, path);
  } else if (fmt != THV_FMT_WCSR && edge_bytes == 4) {
    pread_all(f.fd, rec, 4 * m, base, path);
  } else {
    // interleave (CSR dst + weights -> {dst, w}) or de-interleave (WCSR -> dst)
    const uint64_t CH = 1u << 20;
    std::vector<uint32_t> a(CH), b(2 * CH);
    for (uint64_t i = 0; i < m; i += CH) {
      const uint64_t k = std::min(CH, m - i);
      if (fmt == THV_FMT_WCSR) {
        pread_all(f.fd, b.data(), 8 * k, base + 8 * i, path);
        for (uint64_t j = 0; j < k; j++) rec[i + j] = b[2 * j];
      } else {
        pread_all(f.fd, a.data(), 4 * k, base + 4 * i, path);
        if (file_w) pread_all(f.fd, b.data(), 4 * k, base + 4 * m + 4 * i, path);
        for (uint64_t j = 0; j < k; j++) {
          rec[2 * (i + j)] = a[j];
          rec[2 * (i + j) + 1] = file_w ? b[j] : 20u;  // converter default weight
        }
      }
    }
  }
  check_endpoints(rec, m, edge_bytes / 4, n);
  if (g->csc) {
    std::vector<uint32_t> od(n);
    if (n) pread_all(f.fd, od.data(), 4 * n, base + 4 * m, path);
    g->d_outdeg.alloc(n * 4);
    if (n) HIPC(hipMemcpy(g->d_outdeg.p, od.data(), 4 * n, hipMemcpyHostToDevice));
  }
  graph_build_tables(*g);
  return g.release();
}

Graph *graph_from_arrays(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *dst,
                         const uint32_t *w, const uint32_t *outdeg, int edge_bytes) {
  if (edge_bytes != 4 && edge_bytes != 8) fail(THV_EINVAL, "edge_bytes must be 4 or 8");
  if (n >= 0xFFFFFFFFull) fail(THV_ENOTSUP, "graphs need fewer than 2^32-1 vertices");
  if (!off || (m && !dst)) fail(THV_EINVAL, "null array");
  if (off[n] != m) fail(THV_EFORMAT, "off[n] != m");
  std::unique_ptr<Graph> g(new Graph());
  graph_init_device(*g);
  g->n = n;
  g->m = m;
  g->rec = edge_bytes;
  g->csc = outdeg != nullptr;
  g->h_off.assign(off, off + n + 1);
  check_offsets(g->h_off, n, m);
  g->h_edges.alloc(m * (uint64_t)edge_bytes);
  uint32_t *rec = static_cast<uint32_t *>(g->h_edges.p);
  if (edge_bytes == 4) {
    if (m) memcpy(rec, dst, 4 * m);
  } else {
    for (uint64_t i = 0; i < m; i++) {
      rec[2 * i] = dst[i];
      rec[2 * i + 1] = w ? w[i] : 20u;
    }
  }
  check_endpoints(rec, m, edge_bytes / 4, n);
  if (outdeg) {
    g->d_outdeg.alloc(n * 4);
    if (n) HIPC(hipMemcpy(g->d_outdeg.p, outdeg, 4 * n, hipMemcpyHostToDevice));
  }
  graph_build_tables(*g);
  return g.release();
}

void graph_export(const Graph &g, uint64_t *off, uint32_t *dst, uint32_t *w, uint32_t *outdeg) {
  if (off) memcpy(off, g.h_off.data(), 8 * (g.n + 1));
  const uint32_t *rec = static_cast<const uint32_t *>(g.h_edges.p);
  const int st = g.rec / 4;
  if (dst || w) {
    parallel_for(g.m, [&](uint64_t a, uint64_t b) {
      for (uint64_t i = a; i < b; i++) {
        if (dst) dst[i] = rec[i * st];
        if (w) w[i] = st == 2 ? rec[2 * i + 1] : 20u;
      }
    });
  }
  if (outdeg) {
    if (!g.csc) fail(THV_ENOTSUP, "out-degrees are kept for CSC graphs only");
    HIPC(hipMemcpy(outdeg, g.d_outdeg.p, 4 * g.n, hipMemcpyDeviceToHost));
  }
}

}  // namespace thv
