// raff.cpp -- C ABI of the multi-replica rigid-atom port relaxation
// (include/surfmol_raff.h): device buffers of the shared topology and of the
// per-replica state, uploads and downloads in the reference's float32
// layouts, and the persistent step / relax launches of raff_kernels.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/surfmol_raff.h"
#include "raff_internal.h"

using namespace smr;

#define RCHK(expr)                                                \
    do {                                                          \
        hipError_t _e = (expr);                                   \
        if (_e != hipSuccess) return hip_fail(_e, #expr);         \
    } while (0)

namespace smr {

int raff_seg_for(int n) {
    int s = 16;
    while (s < n) s <<= 1;
    return s;
}

}  // namespace smr

namespace {

hipError_t zero(DevBuf &b, hipStream_t st) { return b.bytes ? hipMemsetAsync(b.ptr, 0, b.bytes, st) : hipSuccess; }

int check_rows(const int *v, int count, int lo, int hi, const char *who, const char *what) {
    for (int i = 0; i < count; i++)
        if (v[i] < lo || v[i] >= hi)
            return fail(SMR_ERR_INVALID, std::string(who) + ": " + what + "[" + std::to_string(i) + "] = " +
                                             std::to_string(v[i]) + " not in [" + std::to_string(lo) + "," +
                                             std::to_string(hi) + ")");
    return SMR_OK;
}

}  // namespace

namespace smr {
int raff_check_cfg(const smr_raff_step_config *c, const char *who) {
    if (!c) return fail(SMR_ERR_INVALID, std::string(who) + ": cfg == NULL");
    if (!(c->dt > 0.f) || !std::isfinite(c->dt) || !std::isfinite(c->relaxation) || !std::isfinite(c->momentum_beta) ||
        !(c->rot_mass_scale >= 0.f) || !std::isfinite(c->rot_mass_scale))
        return fail(SMR_ERR_INVALID, std::string(who) + ": step config out of range (dt=" + std::to_string(c->dt) +
                                         " relaxation=" + std::to_string(c->relaxation) + ")");
    return SMR_OK;
}

int raff_ready(const smr_raff *r, const char *who) {
    if (!r->have_state || !r->have_neighs || !r->have_bk || r->nnode < 0)
        return fail(SMR_ERR_STATE, std::string(who) + ": upload the neighbours, cluster ports, back slots and state "
                                                      "first");
    // back slots address node ports of this replica
    for (int i = 0; i < 4 * r->n; i++)
        if (r->bk_host[i] >= 4 * r->nnode)
            return fail(SMR_ERR_INVALID, std::string(who) + ": bk_slots[" + std::to_string(i) + "] = " +
                                             std::to_string(r->bk_host[i]) + " beyond the " +
                                             std::to_string(4 * r->nnode) + " node ports");
    return SMR_OK;
}

RaffArgs raff_args(smr_raff *r, const smr_raff_step_config *c, int nsteps) {
    RaffArgs A;
    memset(&A, 0, sizeof(A));
    A.natoms = r->n;
    A.nnode = r->nnode;
    A.nsys = r->nsys;
    A.nsteps = nsteps;
    A.radius = (const float *)r->radius.ptr;
    A.neighs = (const int4 *)r->neighs.ptr;
    A.excl1 = (const int4 *)r->excl1.ptr;
    A.excl2 = (const int4 *)r->excl2.ptr;
    A.bk_slots = (const int4 *)r->bk_slots.ptr;
    A.fixmask = (const int *)r->fixmask.ptr;
    A.port_local = (const float4 *)r->port_local.ptr;
    A.kflat = (const float *)r->kflat.ptr;
    A.pos = (float4 *)r->pos.ptr;
    A.quat = (float4 *)r->quat.ptr;
    A.dpos_mom = (float4 *)r->dpos_mom.ptr;
    A.dquat_mom = (float4 *)r->dquat_mom.ptr;
    A.dpos_coll = (float4 *)r->dpos_coll.ptr;
    A.dpos_node = (float4 *)r->dpos_node.ptr;
    A.drot_node = (float4 *)r->drot_node.ptr;
    A.dpos_neigh = (float4 *)r->dpos_neigh.ptr;
    A.metric = (float *)r->metric.ptr;
    A.steps = (int *)r->steps.ptr;
    A.status = (int *)r->status.ptr;
    A.dt = c->dt;
    A.rot_mass_scale = c->rot_mass_scale;
    A.relaxation = c->relaxation;
    A.beta = c->momentum_beta;
    A.collide = r->collide ? 1 : 0;
    return A;
}

}  // namespace smr

namespace {
// float4 rows of a buffer -> 3 floats per row
int download3(const DevBuf &b, size_t row0, size_t nrows, float *out, hipStream_t st) {
    std::vector<float> h(4 * nrows);
    RCHK(hipMemcpyAsync(h.data(), (const float *)b.ptr + 4 * row0, sizeof(float) * h.size(), hipMemcpyDeviceToHost, st));
    RCHK(hipStreamSynchronize(st));
    for (size_t i = 0; i < nrows; i++)
        for (int k = 0; k < 3; k++) out[3 * i + k] = h[4 * i + k];
    return SMR_OK;
}

}  // namespace

extern "C" int smr_raff_create(smr_ctx *ctx, int natoms_per_sys, int nsys, smr_raff **out) {
    if (!ctx || !out) return fail(SMR_ERR_INVALID, "smr_raff_create: NULL argument");
    *out = nullptr;
    if (natoms_per_sys <= 0 || natoms_per_sys > SMR_RAFF_MAX_ATOMS)
        return fail(SMR_ERR_UNSUPPORTED, "smr_raff_create: natoms_per_sys = " + std::to_string(natoms_per_sys) +
                                              " not in [1," + std::to_string(SMR_RAFF_MAX_ATOMS) + "]");
    if (nsys <= 0) return fail(SMR_ERR_INVALID, "smr_raff_create: nsys must be > 0");
    RCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    auto *r = new smr_raff();
    r->ctx = ctx;
    r->n = natoms_per_sys;
    r->nsys = nsys;
    r->seg = raff_seg_for(natoms_per_sys);
    const size_t n = (size_t)natoms_per_sys, na = n * (size_t)nsys;
    hipError_t e = r->radius.alloc(sizeof(float) * n);
    if (e == hipSuccess) e = r->neighs.alloc(sizeof(int4) * n);
    if (e == hipSuccess) e = r->excl1.alloc(sizeof(int4) * n);
    if (e == hipSuccess) e = r->excl2.alloc(sizeof(int4) * n);
    if (e == hipSuccess) e = r->bk_slots.alloc(sizeof(int4) * n);
    if (e == hipSuccess) e = r->fixmask.alloc(sizeof(int) * n);
    if (e == hipSuccess) e = r->pos.alloc(sizeof(float4) * na);
    if (e == hipSuccess) e = r->quat.alloc(sizeof(float4) * na);
    if (e == hipSuccess) e = r->dpos_mom.alloc(sizeof(float4) * na);
    if (e == hipSuccess) e = r->dquat_mom.alloc(sizeof(float4) * na);
    if (e == hipSuccess) e = r->dpos_coll.alloc(sizeof(float4) * na);
    if (e == hipSuccess) e = r->metric.alloc(sizeof(float) * nsys);
    if (e == hipSuccess) e = r->steps.alloc(sizeof(int) * nsys);
    if (e == hipSuccess) e = r->status.alloc(sizeof(int) * nsys);
    if (e == hipSuccess) e = zero(r->radius, st);
    if (e == hipSuccess) e = zero(r->fixmask, st);
    if (e == hipSuccess) e = zero(r->dpos_mom, st);
    if (e == hipSuccess) e = zero(r->dquat_mom, st);
    if (e == hipSuccess) e = zero(r->dpos_coll, st);
    if (e == hipSuccess) e = zero(r->metric, st);
    if (e == hipSuccess) {
        std::vector<float4> id(na, make_float4(0.f, 0.f, 0.f, 1.f));
        e = hipMemcpyAsync(r->quat.ptr, id.data(), sizeof(float4) * na, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess) {
        delete r;
        return hip_fail(e, "smr_raff_create");
    }
    ctx_retain(ctx);
    *out = r;
    return SMR_OK;
}

extern "C" int smr_raff_destroy(smr_raff *r) {
    if (!r) return SMR_OK;
    smr_ctx *c = r->ctx;
    (void)hipSetDevice(c->device);
    delete r;
    ctx_release(c);
    return SMR_OK;
}

extern "C" int smr_raff_upload_radius(smr_raff *r, const float *radius) {
    if (!r || !radius) return fail(SMR_ERR_INVALID, "smr_raff_upload_radius: NULL argument");
    bool any = false;
    for (int i = 0; i < r->n; i++) {
        if (!std::isfinite(radius[i]) || radius[i] < 0.f)
            return fail(SMR_ERR_INVALID, "smr_raff_upload_radius: radius[" + std::to_string(i) +
                                             "] must be finite and non-negative, got " + std::to_string(radius[i]));
        any = any || radius[i] > 0.f;
    }
    RCHK(hipSetDevice(r->ctx->device));
    RCHK(hipMemcpy(r->radius.ptr, radius, sizeof(float) * r->n, hipMemcpyHostToDevice));
    r->collide = any;
    if (!any) RCHK(hipMemset(r->dpos_coll.ptr, 0, r->dpos_coll.bytes));
    return SMR_OK;
}

extern "C" int smr_raff_upload_neighs_and_exclusions(smr_raff *r, const int *neighs, const int *excl1,
                                                     const int *excl2) {
    if (!r || !neighs || !excl1 || !excl2)
        return fail(SMR_ERR_INVALID, "smr_raff_upload_neighs_and_exclusions: NULL argument");
    const char *who = "smr_raff_upload_neighs_and_exclusions";
    int rc = check_rows(neighs, 4 * r->n, -1, r->n, who, "neighs");
    if (!rc) rc = check_rows(excl1, 4 * r->n, -1, r->n, who, "excl1");
    if (!rc) rc = check_rows(excl2, 4 * r->n, -1, r->n, who, "excl2");
    if (rc) return rc;
    RCHK(hipSetDevice(r->ctx->device));
    RCHK(hipMemcpy(r->neighs.ptr, neighs, sizeof(int4) * r->n, hipMemcpyHostToDevice));
    RCHK(hipMemcpy(r->excl1.ptr, excl1, sizeof(int4) * r->n, hipMemcpyHostToDevice));
    RCHK(hipMemcpy(r->excl2.ptr, excl2, sizeof(int4) * r->n, hipMemcpyHostToDevice));
    r->neigh_host.assign(neighs, neighs + 4 * r->n);
    r->have_neighs = true;
    return SMR_OK;
}

extern "C" int smr_raff_upload_fixmask(smr_raff *r, const int *fixmask) {
    if (!r || !fixmask) return fail(SMR_ERR_INVALID, "smr_raff_upload_fixmask: NULL argument");
    RCHK(hipSetDevice(r->ctx->device));
    RCHK(hipMemcpy(r->fixmask.ptr, fixmask, sizeof(int) * r->n, hipMemcpyHostToDevice));
    return SMR_OK;
}

extern "C" int smr_raff_upload_cluster_ports(smr_raff *r, const float *port_local_atoms, const float *k_atoms,
                                             int nnode_per_group) {
    if (!r || !port_local_atoms || !k_atoms) return fail(SMR_ERR_INVALID, "smr_raff_upload_cluster_ports: NULL argument");
    if (nnode_per_group < 0 || nnode_per_group > r->n)
        return fail(SMR_ERR_INVALID, "smr_raff_upload_cluster_ports: nnode_per_group = " +
                                         std::to_string(nnode_per_group) + " not in [0," + std::to_string(r->n) + "]");
    for (int i = 0; i < 16 * nnode_per_group; i++)
        if (!std::isfinite(port_local_atoms[i]))
            return fail(SMR_ERR_INVALID, "smr_raff_upload_cluster_ports: port_local_atoms[" + std::to_string(i) +
                                             "] is not finite");
    for (int i = 0; i < 4 * nnode_per_group; i++)
        if (!std::isfinite(k_atoms[i]))
            return fail(SMR_ERR_INVALID, "smr_raff_upload_cluster_ports: k_atoms[" + std::to_string(i) + "] is not finite");
    RCHK(hipSetDevice(r->ctx->device));
    hipStream_t st = r->ctx->stream;
    const size_t nn = (size_t)std::max(1, nnode_per_group), ns = (size_t)r->nsys;
    hipError_t e = r->port_local.alloc(sizeof(float4) * 4 * nn);
    if (e == hipSuccess) e = r->kflat.alloc(sizeof(float) * 4 * nn);
    if (e == hipSuccess) e = r->dpos_node.alloc(sizeof(float4) * ns * nn);
    if (e == hipSuccess) e = r->drot_node.alloc(sizeof(float4) * ns * nn);
    if (e == hipSuccess) e = r->dpos_neigh.alloc(sizeof(float4) * 4 * ns * nn);
    if (e == hipSuccess) e = zero(r->dpos_node, st);
    if (e == hipSuccess) e = zero(r->drot_node, st);
    if (e == hipSuccess) e = zero(r->dpos_neigh, st);
    // the node rows (atoms 0..nnode-1) of the per-atom arrays
    if (e == hipSuccess && nnode_per_group > 0)
        e = hipMemcpyAsync(r->port_local.ptr, port_local_atoms, sizeof(float4) * 4 * nnode_per_group,
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nnode_per_group > 0)
        e = hipMemcpyAsync(r->kflat.ptr, k_atoms, sizeof(float) * 4 * nnode_per_group, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "smr_raff_upload_cluster_ports");
    r->nnode = nnode_per_group;
    return SMR_OK;
}

extern "C" int smr_raff_upload_bk_slots(smr_raff *r, const int *bk_slots) {
    if (!r || !bk_slots) return fail(SMR_ERR_INVALID, "smr_raff_upload_bk_slots: NULL argument");
    int rc = check_rows(bk_slots, 4 * r->n, -1, 4 * r->n, "smr_raff_upload_bk_slots", "bk_slots");
    if (rc) return rc;
    RCHK(hipSetDevice(r->ctx->device));
    RCHK(hipMemcpy(r->bk_slots.ptr, bk_slots, sizeof(int4) * r->n, hipMemcpyHostToDevice));
    r->bk_host.assign(bk_slots, bk_slots + 4 * r->n);
    r->have_bk = true;
    return SMR_OK;
}

extern "C" int smr_raff_upload_state(smr_raff *r, const float *pos3, const float *inv_mass, const float *quat) {
    if (!r || !pos3 || !inv_mass) return fail(SMR_ERR_INVALID, "smr_raff_upload_state: NULL argument");
    const size_t n = (size_t)r->n, na = n * (size_t)r->nsys;
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(inv_mass[i]) || inv_mass[i] < 0.f)
            return fail(SMR_ERR_INVALID, "smr_raff_upload_state: inv_mass[" + std::to_string(i) + "] = " +
                                             std::to_string(inv_mass[i]));
    std::vector<float4> p(na), q(na);
    const float nan = std::nanf("");
    for (size_t s = 0; s < (size_t)r->nsys; s++)
        for (size_t i = 0; i < n; i++) {
            const size_t a = s * n + i;
            const float w = inv_mass[i];
            if (w <= 1e-12f) {
                p[a] = make_float4(nan, nan, nan, w);
            } else {
                for (int k = 0; k < 3; k++)
                    if (!std::isfinite(pos3[3 * a + k]))
                        return fail(SMR_ERR_INVALID, "smr_raff_upload_state: position of replica " + std::to_string(s) +
                                                         " atom " + std::to_string(i) + " is not finite");
                p[a] = make_float4(pos3[3 * a], pos3[3 * a + 1], pos3[3 * a + 2], w);
            }
            q[a] = quat ? make_float4(quat[4 * a], quat[4 * a + 1], quat[4 * a + 2], quat[4 * a + 3])
                        : make_float4(0.f, 0.f, 0.f, 1.f);
        }
    RCHK(hipSetDevice(r->ctx->device));
    hipStream_t st = r->ctx->stream;
    RCHK(hipMemcpyAsync(r->pos.ptr, p.data(), sizeof(float4) * na, hipMemcpyHostToDevice, st));
    RCHK(hipMemcpyAsync(r->quat.ptr, q.data(), sizeof(float4) * na, hipMemcpyHostToDevice, st));
    RCHK(zero(r->dpos_mom, st));
    RCHK(zero(r->dquat_mom, st));
    RCHK(hipStreamSynchronize(st));
    r->have_state = true;
    return SMR_OK;
}

extern "C" int smr_raff_step(smr_raff *r, const smr_raff_step_config *cfg, int nsteps) {
    if (!r) return fail(SMR_ERR_INVALID, "smr_raff_step: NULL argument");
    int rc = raff_check_cfg(cfg, "smr_raff_step");
    if (!rc) rc = raff_ready(r, "smr_raff_step");
    if (rc) return rc;
    if (nsteps <= 0) return fail(SMR_ERR_INVALID, "smr_raff_step: nsteps must be > 0");
    RCHK(hipSetDevice(r->ctx->device));
    hipStream_t st = r->ctx->stream;
    RaffArgs A = raff_args(r, cfg, nsteps);
    RCHK(launch_raff(A, r->seg, st));
    RCHK(hipStreamSynchronize(st));
    return SMR_OK;
}

extern "C" int smr_raff_relax(smr_raff *r, const smr_raff_step_config *cfg, int max_steps, float tol, int *steps,
                              int *status, float *metric) {
    if (!r) return fail(SMR_ERR_INVALID, "smr_raff_relax: NULL argument");
    int rc = raff_check_cfg(cfg, "smr_raff_relax");
    if (!rc) rc = raff_ready(r, "smr_raff_relax");
    if (rc) return rc;
    if (max_steps <= 0 || !(tol > 0.f)) return fail(SMR_ERR_INVALID, "smr_raff_relax: need max_steps > 0, tol > 0");
    RCHK(hipSetDevice(r->ctx->device));
    hipStream_t st = r->ctx->stream;
    RCHK(zero(r->steps, st));
    RCHK(zero(r->status, st));   // SMR_RAFF_RUNNING
    RaffArgs A = raff_args(r, cfg, max_steps);
    A.relax = 1;
    A.max_steps = max_steps;
    A.tol = tol;
    RCHK(launch_raff(A, r->seg, st));
    if (steps) RCHK(hipMemcpyAsync(steps, r->steps.ptr, sizeof(int) * r->nsys, hipMemcpyDeviceToHost, st));
    if (status) RCHK(hipMemcpyAsync(status, r->status.ptr, sizeof(int) * r->nsys, hipMemcpyDeviceToHost, st));
    if (metric) RCHK(hipMemcpyAsync(metric, r->metric.ptr, sizeof(float) * r->nsys, hipMemcpyDeviceToHost, st));
    RCHK(hipStreamSynchronize(st));
    return SMR_OK;
}

extern "C" int smr_raff_download_pos(const smr_raff *r, float *pos3) {
    if (!r || !pos3) return fail(SMR_ERR_INVALID, "smr_raff_download_pos: NULL argument");
    RCHK(hipSetDevice(r->ctx->device));
    return download3(r->pos, 0, (size_t)r->n * r->nsys, pos3, r->ctx->stream);
}

static int check_isys(const smr_raff *r, int isys, const void *p, const char *who) {
    if (!r || !p) return fail(SMR_ERR_INVALID, std::string(who) + ": NULL argument");
    if (isys < 0 || isys >= r->nsys)
        return fail(SMR_ERR_INVALID, std::string(who) + ": isys = " + std::to_string(isys) + " not in [0," +
                                         std::to_string(r->nsys) + ")");
    return SMR_OK;
}

extern "C" int smr_raff_download_pos_replica(const smr_raff *r, int isys, float *pos3) {
    int rc = check_isys(r, isys, pos3, "smr_raff_download_pos_replica");
    if (rc) return rc;
    RCHK(hipSetDevice(r->ctx->device));
    return download3(r->pos, (size_t)isys * r->n, r->n, pos3, r->ctx->stream);
}

extern "C" int smr_raff_download_quat_replica(const smr_raff *r, int isys, float *quat4) {
    int rc = check_isys(r, isys, quat4, "smr_raff_download_quat_replica");
    if (rc) return rc;
    RCHK(hipSetDevice(r->ctx->device));
    RCHK(hipMemcpy(quat4, (const float4 *)r->quat.ptr + (size_t)isys * r->n, sizeof(float4) * r->n,
                   hipMemcpyDeviceToHost));
    return SMR_OK;
}

extern "C" int smr_raff_download_dpos_coll_replica(const smr_raff *r, int isys, float *d3) {
    int rc = check_isys(r, isys, d3, "smr_raff_download_dpos_coll_replica");
    if (rc) return rc;
    RCHK(hipSetDevice(r->ctx->device));
    return download3(r->dpos_coll, (size_t)isys * r->n, r->n, d3, r->ctx->stream);
}

extern "C" int smr_raff_download_dpos_node_replica(const smr_raff *r, int isys, float *d3) {
    int rc = check_isys(r, isys, d3, "smr_raff_download_dpos_node_replica");
    if (rc) return rc;
    if (r->nnode < 0) return fail(SMR_ERR_STATE, "smr_raff_download_dpos_node_replica: no cluster ports uploaded");
    if (r->nnode == 0) return SMR_OK;
    RCHK(hipSetDevice(r->ctx->device));
    return download3(r->dpos_node, (size_t)isys * r->nnode, r->nnode, d3, r->ctx->stream);
}

extern "C" int smr_raff_download_metric(const smr_raff *r, float *metric) {
    if (!r || !metric) return fail(SMR_ERR_INVALID, "smr_raff_download_metric: NULL argument");
    RCHK(hipSetDevice(r->ctx->device));
    RCHK(hipMemcpy(metric, r->metric.ptr, sizeof(float) * r->nsys, hipMemcpyDeviceToHost));
    return SMR_OK;
}
