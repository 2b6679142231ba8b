// Serving loop: placeholder entry points until the batching server lands.
#include "dte_internal.h"

extern "C" {
int dte_server_create(int32_t, int64_t, int32_t, dte_server**) { dte::set_error(DTE_ERR_INTERNAL, "server not built"); return DTE_ERR_INTERNAL; }
int dte_server_load_forest(dte_server*, const dte_forest*, int32_t*) { return DTE_ERR_INTERNAL; }
int dte_server_submit(dte_server*, int32_t, const float*, int64_t, uint64_t*) { return DTE_ERR_INTERNAL; }
int dte_server_poll(dte_server*, uint64_t, float*, int32_t*) { return DTE_ERR_INTERNAL; }
int dte_server_wait(dte_server*, uint64_t, float*) { return DTE_ERR_INTERNAL; }
int dte_server_get_stats(dte_server*, dte_server_stats*) { return DTE_ERR_INTERNAL; }
int dte_server_destroy(dte_server*) { return DTE_ERR_INTERNAL; }
}
