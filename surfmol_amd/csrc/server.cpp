// server.cpp -- continuous-batching relax server (placeholder until the
// serving loop lands; every entry point reports SMR_ERR_UNSUPPORTED).
#include "../../include/surfmol_relax.h"
#include "smr_internal.h"

using namespace smr;

extern "C" int smr_server_create(smr_ctx *, const smr_gridff *, const smr_server_config *, smr_server **out) {
    if (out) *out = nullptr;
    return fail(SMR_ERR_UNSUPPORTED, "smr_server_create: serving loop not built yet");
}
extern "C" int smr_server_submit(smr_server *, const smr_topology *, const double *, long long *) {
    return fail(SMR_ERR_UNSUPPORTED, "smr_server_submit: serving loop not built yet");
}
extern "C" int smr_server_poll(smr_server *, long long, int *, smr_relax_result *, double *, double *, double *) {
    return fail(SMR_ERR_UNSUPPORTED, "smr_server_poll: serving loop not built yet");
}
extern "C" int smr_server_drain(smr_server *) { return fail(SMR_ERR_UNSUPPORTED, "smr_server_drain: not built"); }
extern "C" int smr_server_destroy(smr_server *) { return SMR_OK; }
