// Serving loop (continuous batching) -- filled in by the serving milestone.
#include "engine.hpp"

extern "C" {
int cof_server_start(cof_ctx *, const cof_server_cfg *, cof_server **out) {
    if (out) *out = nullptr;
    return COF_E_UNSUPPORTED;
}
int cof_server_submit(cof_server *, cof_query *, uint64_t *) { return COF_E_UNSUPPORTED; }
int cof_server_poll(cof_server *, uint64_t, int, cof_result **, double *) { return COF_E_UNSUPPORTED; }
int cof_server_stop(cof_server *) { return COF_E_UNSUPPORTED; }
}
