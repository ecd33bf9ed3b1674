// serve.cpp -- continuous-batching serving loop (placeholder entry points).
#include "rm_internal.h"

extern "C" {
int rm_server_create(rm_ctx *, const rm_index *, int, int, int, int, int, int, int, int64_t, int,
                     int64_t, rm_server **) {
    return rm::fail(RM_ERR_STATE, "serving loop not available in this build");
}
int rm_server_submit(rm_server *, const uint32_t *, int64_t, int64_t *) {
    return rm::fail(RM_ERR_STATE, "serving loop not available in this build");
}
int rm_server_wait(rm_server *, int64_t, int, float *, int32_t *, double *, double *, double *) {
    return rm::fail(RM_ERR_STATE, "serving loop not available in this build");
}
int rm_server_cancel(rm_server *, int64_t) {
    return rm::fail(RM_ERR_STATE, "serving loop not available in this build");
}
int rm_server_stats(rm_server *, int64_t *, int64_t *, int64_t *) {
    return rm::fail(RM_ERR_STATE, "serving loop not available in this build");
}
int rm_server_destroy(rm_server *) { return RM_OK; }
}
