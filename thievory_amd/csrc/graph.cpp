// graph.cpp -- graph loading, the split edge store and the per-graph
// workspace.
//
// Reference counterparts: CSR<uint32>::ReadInputFile (src/graph.cuh:641-815),
// SetPartitionsConfig (graph.cuh:270-304), SetNumStaticEdges
// (graph.cuh:306-341) and the allocations of InitData (graph.cuh:344-638).
// Differences by design: the whole edge array lives in one hipHostMalloc'd
// mapped buffer (no NUMA replicas), the HBM share is chosen by partition
// average degree under a byte budget instead of as a vertex prefix of 90%
// of free memory, and it is filled once at configuration time instead of
// lazily by the filter path.
#include <algorithm>
#include <cstdio>
#include <numeric>

#include "internal.hpp"

namespace thv {

static thread_local std::string g_last_error;

void set_last_error(const std::string &m) { g_last_error = m; }
const char *last_error() { return g_last_error.c_str(); }

[[noreturn]] void fail(int code, const std::string &msg) { throw Error(code, msg); }

void hip_fail(hipError_t e, const char *what, const char *file, int line) {
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e,
           hipGetErrorString(e), what, file, line);
  int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? THV_ENOMEM
                                                                          : THV_EHIP;
  throw Error(code, buf);
}

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

void DevMem::alloc(size_t b) {
  release();
  if (b == 0) b = 8;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    fail(THV_ENOMEM, "hipMalloc of " + std::to_string(b) + " bytes failed");
  }
  bytes = b;
}

void DevMem::release() {
  if (p && own) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  own = true;
}

// Non-coherent (coarse-grained) mapped memory lets the GPU cache the
// read-only edge records in L2; THV_HOST_COHERENT=1 selects fine-grained.
static unsigned host_flags() {
  const char *e = getenv("THV_HOST_COHERENT");
  if (e && e[0] == '1') return hipHostMallocMapped | hipHostMallocCoherent;
  return hipHostMallocMapped | hipHostMallocNonCoherent;
}

void HostPinned::alloc(size_t b) {
  release();
  if (b == 0) b = 8;
  hipError_t e = hipHostMalloc(&p, b, host_flags());
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    fail(THV_ENOMEM, "hipHostMalloc of " + std::to_string(b) + " bytes failed");
  }
  HIPC(hipHostGetDevicePointer(&dev, p, 0));
  bytes = b;
}

void HostPinned::release() {
  if (p) (void)hipHostFree(p);
  p = dev = nullptr;
  bytes = 0;
}

Workspace::~Workspace() {
  if (h_ctr) (void)hipHostFree(h_ctr);
  if (h_aux) (void)hipHostFree(h_aux);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
}

Graph::~Graph() {
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
}

void graph_init_device(Graph &g) {
  HIPC(hipGetDevice(&g.device));
  hipDeviceProp_t prop;
  HIPC(hipGetDeviceProperties(&prop, g.device));
  g.cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  HIPC(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
}

// Device degree tables: degrees and the degree-0 bitmap.
void graph_build_tables(Graph &g) {
  std::vector<uint32_t> deg(vertex_pad(g.n), 0u), iso(bitmap_words(g.n), 0u);
  for (uint64_t v = 0; v < g.n; v++) {
    uint64_t d = g.h_off[v + 1] - g.h_off[v];
    if (d > 0xFFFFFFFFull) fail(THV_ENOTSUP, "vertex degree exceeds 2^32-1");
    deg[v] = (uint32_t)d;
    if (!d) iso[v >> 5] |= 1u << (v & 31);
  }
  g.d_deg.alloc(deg.size() * 4);
  if (g.n) HIPC(hipMemcpy(g.d_deg.p, deg.data(), deg.size() * 4, hipMemcpyHostToDevice));
  g.d_iso.alloc(iso.size() * 4);
  HIPC(hipMemcpy(g.d_iso.p, iso.data(), iso.size() * 4, hipMemcpyHostToDevice));
}

// Default store, built on first use when the caller configured none: like
// CSR::SetNumStaticEdges (graph.cuh:306-341) the edge budget is 90% of the
// HBM still free after the per-vertex tables and a workspace reserve, but
// filled by partition density rather than as a vertex prefix.
void store_ensure(Graph &g) {
  if (g.store.configured) return;
  size_t fr = 0, tot = 0;
  HIPC(hipMemGetInfo(&fr, &tot));
  const uint64_t reserve = g.n * 96 + (g.m / 256 + 1) * 16 + (256ull << 20);
  const uint64_t avail = fr > reserve ? (uint64_t)((fr - reserve) * 0.9) : 0;
  const uint64_t all = g.m * (uint64_t)g.rec;
  store_configure(g, avail >= all ? UINT64_MAX : avail, 1u << 16);
}

// Greedy contiguous vertex ranges of at most part_edges edges; a vertex
// larger than that gets a partition of its own (graph.cuh:270-304).
static void make_partitions(Graph &g, uint32_t part_edges, std::vector<uint32_t> &pv) {
  pv.clear();
  pv.push_back(0);
  uint64_t cur = 0;
  for (uint64_t v = 0; v < g.n; v++) {
    uint64_t d = g.h_off[v + 1] - g.h_off[v];
    if (cur + d <= part_edges || v == pv.back()) {
      cur += d;
    } else {
      pv.push_back((uint32_t)v);
      cur = d;
    }
  }
  pv.push_back((uint32_t)g.n);
  if (g.n == 0) pv.assign({0, 0});
}

// Parent hints pay off wherever bottom-up BFS levels run: a symmetric graph
// with direction-optimising BFS.  A resident store keeps them beside the
// lists (n x 4 bytes more HBM).  A split store takes them from the budget,
// which must leave at least as much again for partitions, and switches to
// the bottom-up layout (sparsest partitions first); only with 4-byte
// records (8-byte {dst, weight} records are loaded for SSSP, whose top-down
// levels want the dense partitions in HBM).
bool store_wants_hints(const Graph &g, uint64_t budget) {
  if (!(g.sym && g.dir_opt && g.parent_hints && g.n > 0)) return false;
  if (budget >= g.m * (uint64_t)g.rec) return true;  // resident
  return g.rec == 4 && budget >= 2 * 4 * g.n;
}

bool store_sparse_first(const Graph &g, bool hints) {
  return g.store_layout == 2 || (g.store_layout == 0 && hints);
}

bool store_stale(const Graph &g) {
  const Store &s = g.store;
  if (!s.configured) return false;
  const bool hints = store_wants_hints(g, s.budget);
  return hints != s.hint_reserved || store_sparse_first(g, hints) != s.sparse_first;
}

// Options that decide the hints or the layout changed: rebuild with the
// same budget.
void store_refresh(Graph &g) {
  Store &s = g.store;
  if (store_stale(g)) store_configure(g, s.budget, s.part_edges);
}

void store_configure(Graph &g, uint64_t budget, uint32_t part_edges) {
  if (part_edges == 0) fail(THV_EINVAL, "partition_edges must be > 0");
  if (g.n >= 0xFFFFFFFFull) fail(THV_ENOTSUP, "graphs need fewer than 2^32-1 vertices");
  Store &s = g.store;
  HIPC(hipStreamSynchronize(g.stream));
  s.configured = false;
  s.part_edges = part_edges;
  s.budget = budget;
  s.hint_reserved = store_wants_hints(g, budget);
  s.d_hint.release();
  // a split store's hints come out of its budget
  if (s.hint_reserved && budget < g.m * (uint64_t)g.rec) budget -= 4 * g.n;
  make_partitions(g, part_edges, s.pv);
  s.P = (uint32_t)(s.pv.size() - 1);
  const uint64_t rec = (uint64_t)g.rec;

  // Residency: partitions ranked by average degree (edges per vertex), taken
  // whole while the byte budget allows.  Densest first: the hubs, which
  // top-down levels and repeated relaxations read most, stay in HBM.
  // Sparsest first under the bottom-up layout: a bottom-up level pays one
  // host request per unvisited vertex with a host list, so HBM holds the
  // most vertices per byte, and the hubs, whose lists are read top-down
  // only from small frontiers or never (hints), stay on the host.
  // (Sparsest first counts only vertices with edges: isolated vertices
  // cost no request.)
  const bool sparse_first = store_sparse_first(g, s.hint_reserved);
  s.sparse_first = sparse_first;
  std::vector<uint32_t> order(s.P), pverts(s.P);
  std::iota(order.begin(), order.end(), 0u);
  auto pedges = [&](uint32_t p) { return g.h_off[s.pv[p + 1]] - g.h_off[s.pv[p]]; };
  for (uint32_t p = 0; p < s.P; p++) {
    uint32_t c = 0;
    if (sparse_first) {
      for (uint64_t v = s.pv[p]; v < s.pv[p + 1]; v++) c += g.h_off[v + 1] != g.h_off[v];
    } else {
      c = s.pv[p + 1] - s.pv[p];
    }
    pverts[p] = c ? c : 1;
  }
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    uint64_t ea = pedges(a), eb = pedges(b);
    uint64_t va = pverts[a], vb = pverts[b];
    // ea/va > eb/vb  <=>  ea*vb > eb*va  (128-bit products)
    unsigned __int128 l = (unsigned __int128)ea * vb, r = (unsigned __int128)eb * va;
    return sparse_first ? l < r : l > r;
  });
  s.pres.assign(s.P, 0);
  uint64_t used = 0;
  s.dev_edges = 0;
  s.Pdev = 0;
  std::vector<uint32_t> chosen;  // partitions taken, in the order taken
  for (uint32_t p : order) {
    uint64_t e = pedges(p);
    if (e == 0) {  // nothing to store; count it as device resident
      s.pres[p] = 1;
      s.Pdev++;
      continue;
    }
    if (budget != UINT64_MAX && used + e * rec > budget) continue;
    s.pres[p] = 1;
    used += e * rec;
    s.dev_edges += e;
    s.Pdev++;
    chosen.push_back(p);
  }

  // Short-list side area (bottom-up layout): the lists of at most
  // SIDE_DEGREE records that are left in host partitions are copied to HBM
  // as well, so a bottom-up level reads every short list from HBM.  They
  // are few (the sparsest partitions are resident already) and charged to
  // the budget: the last partitions taken make room if needed.
  constexpr uint32_t SIDE_DEGREE = 16;
  auto short_edges = [&](uint32_t p) {
    uint64_t e = 0;
    for (uint64_t v = s.pv[p]; v < s.pv[p + 1]; v++) {
      const uint64_t d = g.h_off[v + 1] - g.h_off[v];
      if (d >= 1 && d <= SIDE_DEGREE) e += d;
    }
    return e;
  };
  uint64_t side_edges = 0;
  if (sparse_first && budget != UINT64_MAX) {
    for (uint32_t p = 0; p < s.P; p++)
      if (!s.pres[p]) side_edges += short_edges(p);
    // how many of the last partitions taken must make room; if the side
    // area does not fit even with none of them, keep them all and go
    // without the side area
    uint64_t u = used, se = side_edges;
    size_t evict = 0;
    while (u + se * rec > budget && evict < chosen.size()) {
      const uint32_t p = chosen[chosen.size() - 1 - evict];
      u -= pedges(p) * rec;
      se += short_edges(p);
      evict++;
    }
    if (u + se * rec > budget) {
      side_edges = 0;  // no room at all: the partitions stay
    } else {
      for (size_t i = 0; i < evict; i++) {
        const uint32_t p = chosen.back();
        chosen.pop_back();
        s.pres[p] = 0;
        s.dev_edges -= pedges(p);
        s.Pdev--;
      }
      used = u;
      side_edges = se;
    }
  }

  // Slab: resident partitions packed in vertex order; contiguous runs of
  // resident partitions are copied with one transfer each.
  s.slab.alloc(s.dev_edges * rec);
  std::vector<uint64_t> slab_off(s.P, 0);
  uint64_t pos = 0;
  for (uint32_t p = 0; p < s.P;) {
    if (!s.pres[p] || pedges(p) == 0) {
      p++;
      continue;
    }
    uint32_t q = p;
    uint64_t run_start_edge = g.h_off[s.pv[p]];
    uint64_t run_pos = pos;
    while (q < s.P && s.pres[q]) {
      slab_off[q] = pos;
      pos += pedges(q);
      q++;
    }
    uint64_t run_edges = pos - run_pos;
    if (run_edges)
      HIPC(hipMemcpy(s.slab.as<char>() + run_pos * rec,
                     (const char *)g.h_edges.p + run_start_edge * rec, run_edges * rec,
                     hipMemcpyHostToDevice));
    p = q;
  }

  // Side area: the short host lists, packed in vertex order.
  s.side.alloc(side_edges * rec);
  s.side_edges = side_edges;
  s.dev_edges += side_edges;
  if (side_edges) {
    std::vector<char> stage(side_edges * rec);
    uint64_t at = 0;
    for (uint32_t p = 0; p < s.P; p++) {
      if (s.pres[p]) continue;
      for (uint64_t v = s.pv[p]; v < s.pv[p + 1]; v++) {
        const uint64_t d = g.h_off[v + 1] - g.h_off[v];
        if (d < 1 || d > SIDE_DEGREE) continue;
        memcpy(stage.data() + at * rec, (const char *)g.h_edges.p + g.h_off[v] * rec, d * rec);
        at += d;
      }
    }
    HIPC(hipMemcpy(s.side.p, stage.data(), side_edges * rec, hipMemcpyHostToDevice));
  }

  // Per-vertex first-record address, residency bit and partition id.
  std::vector<uint64_t> eptr(g.n);
  std::vector<uint32_t> rbit(bitmap_words(g.n), 0u), vpart(g.n);
  const uint64_t slab_base = (uint64_t)(uintptr_t)s.slab.p;
  const uint64_t side_base = (uint64_t)(uintptr_t)s.side.p;
  const uint64_t host_base = (uint64_t)(uintptr_t)g.h_edges.dev;
  s.host_vertices = s.host_short_lists = s.host_short_edges = 0;
  uint64_t side_at = 0;
  for (uint32_t p = 0; p < s.P; p++) {
    uint64_t pe0 = g.h_off[s.pv[p]];
    for (uint64_t v = s.pv[p]; v < s.pv[p + 1]; v++) {
      vpart[v] = p;
      const uint64_t d = g.h_off[v + 1] - g.h_off[v];
      if (s.pres[p]) {
        eptr[v] = slab_base + (slab_off[p] + (g.h_off[v] - pe0)) * rec;
        rbit[v >> 5] |= 1u << (v & 31);
      } else if (side_edges && d >= 1 && d <= SIDE_DEGREE) {
        eptr[v] = side_base + side_at * rec;
        rbit[v >> 5] |= 1u << (v & 31);
        side_at += d;
      } else {
        eptr[v] = host_base + g.h_off[v] * rec;
        s.host_vertices += d != 0;
        if (d >= 1 && d <= SIDE_DEGREE) {
          s.host_short_lists++;
          s.host_short_edges += d;
        }
      }
    }
  }
  s.d_eptr.alloc(g.n * 8);
  if (g.n) HIPC(hipMemcpy(s.d_eptr.p, eptr.data(), g.n * 8, hipMemcpyHostToDevice));
  s.d_rbit.alloc(rbit.size() * 4);
  HIPC(hipMemcpy(s.d_rbit.p, rbit.data(), rbit.size() * 4, hipMemcpyHostToDevice));
  s.d_vpart.alloc(g.n * 4);
  if (g.n) HIPC(hipMemcpy(s.d_vpart.p, vpart.data(), g.n * 4, hipMemcpyHostToDevice));
  s.d_pv.alloc((s.P + 1) * 4);
  HIPC(hipMemcpy(s.d_pv.p, s.pv.data(), (s.P + 1) * 4, hipMemcpyHostToDevice));
  s.d_pres.alloc(s.P ? s.P : 1);
  if (s.P) HIPC(hipMemcpy(s.d_pres.p, s.pres.data(), s.P, hipMemcpyHostToDevice));
  s.configured = true;
  if (s.hint_reserved) build_parent_hints(g);
}

// PageRank push hubs: count the destinations of every 64th record (host
// copy of the edge array), keep the PR_HUBS most frequent ones that were
// seen at least 4 times (256 records expected), sorted.
void ensure_pr_hubs(Graph &g) {
  if (g.hubs_built) return;
  std::vector<uint32_t> cnt(g.n, 0);
  const char *rec = static_cast<const char *>(g.h_edges.p);
  for (uint64_t e = 0; e < g.m; e += 64) {
    uint32_t w;
    memcpy(&w, rec + e * g.rec, 4);
    if (w < g.n) cnt[w]++;
  }
  std::vector<uint32_t> cand;
  for (uint64_t v = 0; v < g.n; v++)
    if (cnt[v] >= 4) cand.push_back((uint32_t)v);
  if (cand.size() > PR_HUBS) {
    std::nth_element(cand.begin(), cand.begin() + PR_HUBS, cand.end(),
                     [&](uint32_t a, uint32_t b) { return cnt[a] > cnt[b]; });
    cand.resize(PR_HUBS);
  }
  std::sort(cand.begin(), cand.end());
  std::vector<uint32_t> bits(bitmap_words(g.n), 0u);
  for (uint32_t v : cand) bits[v >> 5] |= 1u << (v & 31);
  g.d_hubs.alloc(std::max<size_t>(1, cand.size()) * 4);
  if (!cand.empty())
    HIPC(hipMemcpy(g.d_hubs.p, cand.data(), cand.size() * 4, hipMemcpyHostToDevice));
  g.d_hubbit.alloc(bits.size() * 4);
  HIPC(hipMemcpy(g.d_hubbit.p, bits.data(), bits.size() * 4, hipMemcpyHostToDevice));
  g.nhubs = (uint32_t)cand.size();
  g.hubs_built = true;
}

void QueueSet::ensure(const Graph &g, uint32_t chunk) {
  if (P != g.store.P || !ptouch.p) {
    ptouch.alloc((uint64_t)(g.store.P + 1) * 4);
    HIPC(hipMemset(ptouch.p, 0, (uint64_t)(g.store.P + 1) * 4));
    stamp = 0;
    P = g.store.P;
  }
  const uint64_t tiles = g.n / TILE + 1;
  if (slist.bytes < tiles * TILE * 4) {
    slist.alloc(tiles * TILE * 4);
    scnt.alloc(tiles * 4 * 4);
    llist.alloc(tiles * TILE * 4);
    lcnt.alloc(tiles * 4 * 4);
    loff.alloc((tiles * 4 + 1) * 4);
    for (int k = 0; k < 2; k++) {
      clist[k].alloc(tiles * TILE * 4);
      ccnt[k].alloc(tiles * 4 * 4);
    }
  }
  for (int k = 0; k < 2; k++) {
    if (tcnt[k].bytes < tiles * 2 * 4) {
      tcnt[k].alloc(tiles * 2 * 4);
      tdeg[k].alloc(tiles * 2 * 8);
      const uint64_t segs = tiles / 64 + 1;
      ssum[k].alloc(segs * 4 * 8);
      HIPC(hipMemset(ssum[k].p, 0, segs * 4 * 8));
      sbase[k].alloc(segs * 4 * 8);
    }
  }
  if (n == g.n && m == g.m && C == chunk && ctr.p) return;
  for (int q = 0; q < 2; q++) {
    qv[q].alloc((g.n + 1) * 4);
    qe[q].alloc((g.n + 1) * 8);
    qc[q].alloc((g.m / chunk + 2) * 4);
  }
  ctr.alloc(CTR_SLOTS * sizeof(LevelCtr));
  n = g.n;
  m = g.m;
  C = chunk;
}

void workspace_ensure(Graph &g) {
  store_ensure(g);
  Workspace &w = g.ws;
  w.qs.ensure(g, 256u * (uint32_t)g.shape.ept);
  if (w.n == g.n && w.m == g.m && w.val.p) return;
  const uint64_t words = bitmap_words(g.n);
  for (auto &f : w.front) f.alloc(words * 4);
  w.vis.alloc(words * 4);
  w.fsum.alloc(fsum_words(g.n) * 4);
  w.aux.alloc(16 * sizeof(unsigned long long));
  w.val.alloc((g.n + 1) * 4);
  // fine-grained: a kernel's stores land in host memory without a cache
  // flush, and the host reads them after the stream sync
  constexpr unsigned HF = hipHostMallocMapped | hipHostMallocCoherent;
  if (!w.h_ctr) {
    HIPC(hipHostMalloc((void **)&w.h_ctr, CTR_SLOTS * sizeof(LevelCtr), HF));
    memset(w.h_ctr, 0, CTR_SLOTS * sizeof(LevelCtr));
    HIPC(hipHostGetDevicePointer((void **)&w.h_ctr_dev, w.h_ctr, 0));
  }
  if (!w.h_aux) {
    HIPC(hipHostMalloc((void **)&w.h_aux, 16 * sizeof(unsigned long long), HF));
    memset(w.h_aux, 0, 16 * sizeof(unsigned long long));
    HIPC(hipHostGetDevicePointer((void **)&w.h_aux_dev, w.h_aux, 0));
  }
  for (auto &e : w.ev)
    if (!e) HIPC(hipEventCreate(&e));
  w.n = g.n;
  w.m = g.m;
}

}  // namespace thv
