// noalloc.hip -- BFS whose per-call buffers live in a static device array.
//
// thv_bfs keeps its workspace (frontier queues, tile tables, bitmaps, the
// value array, level counters) on the graph handle and allocates it on the
// first call.  thv_bfs_noalloc runs the same top-down kernels
// (compaction + edge-balanced expansion) over a workspace carved from
// `scratch`, a file-scope device array of THV_NOALLOC_SCRATCH_BYTES that
// every GPU holds from module load to process end, so the call itself never
// allocates device or pinned memory.  It is for callers that cannot
// allocate (code under stream capture, tight serving loops).
//
// In host code `scratch` names a host-side stand-in, not device memory:
// the device address comes from hipGetSymbolAddress, taken once per device
// with that device current, and only that address is used.  Every call on
// one GPU shares its array and finds the last call's contents, so whatever
// must start at zero is cleared by a fill kernel at the start of the call;
// calls that overlap on one GPU race.
#include <mutex>

#include "internal.hpp"

namespace thv {

static __device__ __align__(256) unsigned char scratch[THV_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// The device address of this device's array (the device is current).
void *scratch_address(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) fail(THV_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      fail(THV_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// Hands out 256-byte-aligned pieces of [base, base + cap); with base null it
// only adds up the sizes.
struct Carver {
  unsigned char *base;
  uint64_t cap, used = 0;
  void take(DevMem &m, uint64_t bytes) {
    const uint64_t b = bytes ? bytes : 8;
    if (base) m.borrow(base + used, b);
    used += div_up(b, 256) * 256;
  }
};

// The workspace of a top-down BFS: what launch_compact and the BFS
// expansion read and write (tile-table set 0 only; no probe lists).
void lay_out(Graph &g, Workspace &w, Carver &c) {
  const uint64_t words = bitmap_words(g.n);
  const uint64_t tiles = g.n / TILE + 1, segs = tiles / 64 + 1;
  const uint32_t C = 256u * (uint32_t)g.shape.ept;
  QueueSet &q = w.qs;
  for (int k = 0; k < 2; k++) {
    c.take(q.qv[k], (g.n + 1) * 4);
    c.take(q.qe[k], (g.n + 1) * 8);
    c.take(q.qc[k], (g.m / C + 2) * 4);
  }
  c.take(q.ctr, CTR_SLOTS * sizeof(LevelCtr));
  c.take(q.tcnt[0], tiles * 2 * 4);
  c.take(q.tdeg[0], tiles * 2 * 8);
  c.take(q.ssum[0], segs * 4 * 8);
  c.take(q.sbase[0], segs * 4 * 8);
  c.take(q.ptouch, ((uint64_t)g.store.P + 1) * 4);
  for (auto &f : w.front) c.take(f, words * 4);
  c.take(w.vis, words * 4);
  c.take(w.aux, 16 * sizeof(unsigned long long));
  c.take(w.val, (g.n + 1) * 4);
  q.n = g.n;
  q.m = g.m;
  q.C = C;
  q.P = g.store.P;
  q.stamp = 0;
  w.n = g.n;
  w.m = g.m;
}

// Restores the caller's current device when the call ends.
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    HIPC(hipGetDevice(&prev));
    if (prev != dev) HIPC(hipSetDevice(dev));
  }
  ~DeviceScope() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace

uint64_t noalloc_bfs_bytes(Graph &g) {
  Workspace w;
  Carver c{nullptr, 0};
  lay_out(g, w, c);
  return c.used;
}

void run_bfs_noalloc(Graph &g, uint32_t src, uint32_t *out, thv_stats *st) {
  // the store must exist already: building it allocates
  if (!g.store.configured || !g.store.d_eptr.p)
    fail(THV_ESTATE, "the graph's edge store is not built: call thv_store_config first");
  const uint64_t need = noalloc_bfs_bytes(g);
  if (need > THV_NOALLOC_SCRATCH_BYTES)
    fail(THV_EINVAL, "graph too large for the static BFS workspace: needs " +
                         std::to_string(need) + " bytes, the array holds " +
                         std::to_string((uint64_t)THV_NOALLOC_SCRATCH_BYTES) +
                         " (about 30 bytes per vertex)");
  DeviceScope ds(g.device);
  unsigned char *base = static_cast<unsigned char *>(scratch_address(g.device));
  Workspace w;  // views of the array; no events, host counters on the stack
  Carver c{base, THV_NOALLOC_SCRATCH_BYTES};
  lay_out(g, w, c);
  LevelCtr h_ctr[2];
  w.h_ctr = h_ctr;
  struct Unhook {  // the host counters are not the workspace's to free
    Workspace &w;
    ~Unhook() { w.h_ctr = nullptr; }
  } unhook{w};
  hipStream_t s = g.stream;
  // the array keeps the last call's contents: the segment sums and the
  // partition stamps must start at zero (fill kernels, no memset nodes)
  const uint64_t segs = (g.n / TILE + 1) / 64 + 1;
  launch_fill_u32(w.qs.ssum[0].as<uint32_t>(), 0u, segs * 4 * 2, s);
  launch_fill_u32(w.qs.ptouch.as<uint32_t>(), 0u, (uint64_t)g.store.P + 1, s);
  bfs_topdown_in(g, w, src, st);
  if (out && g.n) {
    HIPC(hipMemcpyAsync(out, w.val.p, g.n * 4, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
  }
  if (st && out) {
    uint64_t r = 0;
    for (uint64_t v = 0; v < g.n; v++) r += out[v] != (uint32_t)(g.n + 1);
    st->reached = r;
  }
}

}  // namespace thv
