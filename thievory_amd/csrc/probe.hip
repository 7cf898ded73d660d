// probe.hip -- link and memory ceilings of the device (thv_link_probe).
//
// The engine reads host-resident edge records zero-copy; this measures the
// rates that bound it, on the same kind of buffer the store uses
// (hipHostMalloc, mapped, non-coherent):
//   zero-copy   a streaming kernel reading the mapped host buffer
//   dma         hipMemcpyAsync pinned host -> device
//   hbm_read    the same streaming kernel over a device buffer
//   hbm_copy    hipMemcpyAsync device -> device (bytes read + written)
// Each is the best of three timed runs after a warm-up, by HIP events.
#include "internal.hpp"

namespace thv {

namespace {

// Reads n16 16-byte words and folds them; the fold is written only when it
// equals `key` (the caller passes a value the data cannot fold to), so the
// loads are kept.
__global__ void __launch_bounds__(256) k_stream_read(const uint4 *p, uint64_t n16,
                                                     unsigned long long *sink, uint32_t key) {
  uint32_t f = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 x = p[i];
    f ^= x.x ^ x.y ^ x.z ^ x.w;
  }
  if (f == key) atomicAdd(sink, 1ull);
}

}  // namespace

void link_probe(uint64_t bytes, double out[4]) {
  bytes &= ~15ull;
  if (bytes < (1u << 20)) fail(THV_EINVAL, "probe needs at least 1 MiB");
  int dev = 0, cus = 0;
  HIPC(hipGetDevice(&dev));
  HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  hipStream_t s;
  HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{s};
  HostPinned host;
  host.alloc(bytes);
  memset(host.p, 0x5A, bytes);  // folds to 0 per 16 bytes: never the key
  DevMem a, b, sink;
  a.alloc(bytes);
  b.alloc(bytes);
  sink.alloc(8);
  HIPC(hipMemsetAsync(a.p, 0x3C, bytes, s));
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  struct EventGuard {
    hipEvent_t a, b;
    ~EventGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } eg{e0, e1};
  const int grid = 8 * cus;
  auto best = [&](auto op) {
    op();  // warm-up
    float bms = 1e30f;
    for (int r = 0; r < 3; r++) {
      HIPC(hipEventRecord(e0, s));
      op();
      HIPC(hipEventRecord(e1, s));
      HIPC(hipEventSynchronize(e1));
      float ms = 0;
      HIPC(hipEventElapsedTime(&ms, e0, e1));
      bms = std::min(bms, ms);
    }
    return (double)bms * 1e-3;
  };
  const double t_zc = best([&] {
    hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(256), 0, s,
                       static_cast<const uint4 *>(host.dev), bytes / 16,
                       sink.as<unsigned long long>(), 0x9E3779B9u);
  });
  const double t_dma = best([&] { HIPC(hipMemcpyAsync(a.p, host.p, bytes, hipMemcpyHostToDevice, s)); });
  const double t_rd = best([&] {
    hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(256), 0, s,
                       static_cast<const uint4 *>(a.p), bytes / 16,
                       sink.as<unsigned long long>(), 0x9E3779B9u);
  });
  const double t_cp = best([&] { HIPC(hipMemcpyAsync(b.p, a.p, bytes, hipMemcpyDeviceToDevice, s)); });
  HIPC(hipGetLastError());
  out[0] = bytes / t_zc / 1e9;
  out[1] = bytes / t_dma / 1e9;
  out[2] = bytes / t_rd / 1e9;
  out[3] = 2.0 * bytes / t_cp / 1e9;
}

}  // namespace thv
