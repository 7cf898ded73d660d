// batch.hip -- multi-source BFS / SSSP (filled in below).
#include "internal.hpp"

namespace thv {
void run_batch(Graph &, int, const uint32_t *, uint32_t, uint64_t *, uint64_t *, uint32_t *,
               uint32_t *, thv_stats *) {
  fail(THV_ENOTSUP, "batched queries are not built yet");
}
void autotune(Graph &, const char *, int *) { fail(THV_ENOTSUP, "autotuner not built yet"); }
}  // namespace thv
