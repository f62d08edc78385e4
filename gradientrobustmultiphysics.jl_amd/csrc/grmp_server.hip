// Request server (continuous batching) -- filled in by the serving layer.
#include "grmp_engine.h"

struct grmp_server {
    grmp_ctx *ctx;
};

extern "C" int grmp_server_create(grmp_ctx *ctx, const grmp_server_config *, grmp_server **out)
{
    if (out) *out = nullptr;
    return grmp_fail(ctx, GRMP_ERR_UNSUPPORTED, "server not built");
}
extern "C" int grmp_server_submit(grmp_server *, const grmp_request *, int64_t *) { return GRMP_ERR_UNSUPPORTED; }
extern "C" int grmp_server_poll(grmp_server *, int64_t, grmp_result *) { return GRMP_ERR_UNSUPPORTED; }
extern "C" int grmp_server_wait(grmp_server *, int64_t, grmp_result *) { return GRMP_ERR_UNSUPPORTED; }
extern "C" double grmp_server_clock(const grmp_server *) { return 0.0; }
extern "C" int grmp_server_destroy(grmp_server *) { return GRMP_OK; }
extern "C" int64_t grmp_request_footprint(const grmp_request *) { return -1; }
extern "C" int64_t grmp_request_ndofs(const grmp_request *) { return -1; }
