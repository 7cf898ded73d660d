// serve.cpp -- continuous-batching query server (filled in below).
#include "internal.hpp"

extern "C" {
int thv_serve_start(thv_graph *, const thv_serve_cfg *, thv_server **) { return THV_ENOTSUP; }
int thv_serve_submit(thv_server *, int, uint32_t, uint64_t *) { return THV_ENOTSUP; }
int thv_serve_poll(thv_server *, uint64_t, thv_result *, uint32_t *) { return THV_ENOTSUP; }
int thv_serve_drain(thv_server *) { return THV_ENOTSUP; }
void thv_serve_stop(thv_server *) {}
}
