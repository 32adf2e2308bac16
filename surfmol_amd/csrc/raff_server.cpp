// raff_server.cpp -- continuous-batching relax server for rigid-atom port
// replicas (smr_raff_server_* of include/surfmol_raff.h).
//
// The server owns a replica set of `capacity` slots that shares one
// molecule's topology (copied from a template replica set).  submit() queues
// a request (positions and orientations of one replica); a worker thread
// loops:
//   admit queued requests into free slots (FIFO) and upload their state;
//   run one chunk: chunk_steps steps of every resident replica in one
//     persistent launch of the relax kernel, whose per-slot status and step
//     count carry over between chunks (free slots are skipped);
//   read the statuses back; a replica that converged or hit the step cap is
//     copied out, its slot freed and its request marked done.
// Requests that arrive while a chunk runs join at the next chunk.  A
// replica's arithmetic does not depend on its slot, its neighbours or the
// chunk boundaries (its state round-trips exactly through device memory), so
// its result equals smr_raff_relax's on a one-replica set, bit for bit.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "raff_internal.h"

using namespace smr;

#define SMR_RAFF_IDLE 3   // a free slot: the kernel skips it

namespace {
using clk = std::chrono::steady_clock;

struct RaffReq {
    long long id = 0;
    std::vector<float4> pos4, quat;
    double t_submit = 0.0, t_done = 0.0;
    bool done = false;
    int steps = 0, status = SMR_RAFF_RUNNING;
    float metric = 0.f;
    std::vector<float> pos_out;   // [3n]
};
}  // namespace

struct smr_raff_server {
    smr_raff *r = nullptr;        // capacity slots
    hipStream_t stream = nullptr; // the worker's own stream
    smr_raff_server_config cfg;
    std::vector<float> inv_mass;  // [n]
    clk::time_point t0;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<RaffReq *> queue;
    std::unordered_map<long long, RaffReq *> reqs;
    std::vector<long long> slot_id;   // request in each slot, -1 free
    std::vector<int> free_slots;
    std::vector<int> hsteps, hstatus;
    std::vector<float> hmetric;
    long long next_id = 1;
    int resident = 0;
    double submitted = 0, retired = 0, chunks = 0, peak = 0, device_s = 0;
    bool stop = false;
    std::string worker_error;
    std::thread worker;
};

namespace {

double now_s(const smr_raff_server *s) {
    return std::chrono::duration<double>(clk::now() - s->t0).count();
}

// one chunk: admit, launch, retire.  Returns a HIP error or hipSuccess.
hipError_t raff_chunk(smr_raff_server *s) {
    smr_raff *r = s->r;
    const int n = r->n, cap = r->nsys;
    hipStream_t st = s->stream;
    std::vector<std::pair<int, RaffReq *>> adm;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        while (!s->queue.empty() && !s->free_slots.empty()) {
            const int slot = s->free_slots.back();
            s->free_slots.pop_back();
            RaffReq *q = s->queue.front();
            s->queue.pop_front();
            s->slot_id[slot] = q->id;
            adm.emplace_back(slot, q);
            s->resident++;
        }
        s->peak = std::max(s->peak, (double)s->resident);
    }
    hipError_t e = hipSuccess;
    for (auto &a : adm) {
        const size_t off = (size_t)a.first * n;
        if (e == hipSuccess)
            e = hipMemcpyAsync((float4 *)r->pos.ptr + off, a.second->pos4.data(), sizeof(float4) * n,
                               hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync((float4 *)r->quat.ptr + off, a.second->quat.data(), sizeof(float4) * n,
                               hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemsetAsync((float4 *)r->dpos_mom.ptr + off, 0, sizeof(float4) * n, st);
        if (e == hipSuccess) e = hipMemsetAsync((float4 *)r->dquat_mom.ptr + off, 0, sizeof(float4) * n, st);
        s->hstatus[a.first] = SMR_RAFF_RUNNING;
        s->hsteps[a.first] = 0;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(r->status.ptr, s->hstatus.data(), sizeof(int) * cap, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(r->steps.ptr, s->hsteps.data(), sizeof(int) * cap, hipMemcpyHostToDevice, st);
    RaffArgs A = raff_args(r, &s->cfg.step, s->cfg.chunk_steps);
    A.relax = 1;
    A.max_steps = s->cfg.max_steps;
    A.tol = s->cfg.tol;
    const auto tk = clk::now();
    if (e == hipSuccess) e = launch_raff(A, r->seg, st);
    if (e == hipSuccess) e = hipMemcpyAsync(s->hstatus.data(), r->status.ptr, sizeof(int) * cap, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(s->hsteps.data(), r->steps.ptr, sizeof(int) * cap, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(s->hmetric.data(), r->metric.ptr, sizeof(float) * cap, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    s->device_s += std::chrono::duration<double>(clk::now() - tk).count();
    // retire
    std::vector<float4> buf(n);
    std::vector<std::pair<int, RaffReq *>> out;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (int slot = 0; slot < cap; slot++) {
            if (s->slot_id[slot] < 0 || s->hstatus[slot] == SMR_RAFF_RUNNING) continue;
            out.emplace_back(slot, s->reqs[s->slot_id[slot]]);
        }
    }
    for (auto &o : out) {
        e = hipMemcpyAsync(buf.data(), (const float4 *)r->pos.ptr + (size_t)o.first * n, sizeof(float4) * n,
                           hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        RaffReq *q = o.second;
        q->pos_out.resize(3 * (size_t)n);
        for (int i = 0; i < n; i++) {
            q->pos_out[3 * i] = buf[i].x;
            q->pos_out[3 * i + 1] = buf[i].y;
            q->pos_out[3 * i + 2] = buf[i].z;
        }
        q->steps = s->hsteps[o.first];
        q->status = s->hstatus[o.first];
        q->metric = s->hmetric[o.first];
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const double t = now_s(s);
        for (auto &o : out) {
            o.second->t_done = t;
            o.second->done = true;
            s->slot_id[o.first] = -1;
            s->hstatus[o.first] = SMR_RAFF_IDLE;
            s->free_slots.push_back(o.first);
            s->resident--;
            s->retired++;
        }
        s->chunks++;
    }
    if (!out.empty()) s->cv_done.notify_all();
    return hipSuccess;
}

void raff_worker(smr_raff_server *s) {
    (void)hipSetDevice(s->r->ctx->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->queue.empty() || s->resident > 0; });
            if (s->stop) return;
        }
        const hipError_t e = raff_chunk(s);
        if (e != hipSuccess) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->worker_error = std::string("raff server chunk: ") + hipGetErrorString(e);
            s->stop = true;
            s->cv_done.notify_all();
            return;
        }
    }
}

}  // namespace

extern "C" int smr_raff_server_create(const smr_raff *tmpl, const smr_raff_server_config *cfg, smr_raff_server **out) {
    if (!tmpl || !cfg || !out) return fail(SMR_ERR_INVALID, "smr_raff_server_create: NULL argument");
    *out = nullptr;
    int rc = raff_check_cfg(&cfg->step, "smr_raff_server_create");
    if (!rc) rc = raff_ready(tmpl, "smr_raff_server_create (template)");
    if (rc) return rc;
    if (cfg->capacity <= 0 || cfg->chunk_steps <= 0 || cfg->max_steps <= 0 || !(cfg->tol > 0.f))
        return fail(SMR_ERR_INVALID, "smr_raff_server_create: need capacity, chunk_steps, max_steps > 0 and tol > 0");
    const int n = tmpl->n;
    smr_raff *r = nullptr;
    rc = smr_raff_create(tmpl->ctx, n, cfg->capacity, &r);
    if (rc) return rc;
    auto fail_r = [&](hipError_t e) {
        smr_raff_destroy(r);
        return hip_fail(e, "smr_raff_server_create");
    };
    hipStream_t st = tmpl->ctx->stream;
    // the template's shared topology
    r->nnode = tmpl->nnode;
    r->collide = tmpl->collide;
    const size_t nn = (size_t)std::max(1, tmpl->nnode), ns = (size_t)cfg->capacity;
    hipError_t e = r->port_local.alloc(sizeof(float4) * 4 * nn);
    if (e == hipSuccess) e = r->kflat.alloc(sizeof(float) * 4 * nn);
    if (e == hipSuccess) e = r->dpos_node.alloc(sizeof(float4) * ns * nn);
    if (e == hipSuccess) e = r->drot_node.alloc(sizeof(float4) * ns * nn);
    if (e == hipSuccess) e = r->dpos_neigh.alloc(sizeof(float4) * 4 * ns * nn);
    const struct { smr::DevBuf *dst; const smr::DevBuf *src; } cp[] = {
        {&r->radius, &tmpl->radius}, {&r->neighs, &tmpl->neighs}, {&r->excl1, &tmpl->excl1},
        {&r->excl2, &tmpl->excl2}, {&r->bk_slots, &tmpl->bk_slots}, {&r->fixmask, &tmpl->fixmask},
        {&r->port_local, &tmpl->port_local}, {&r->kflat, &tmpl->kflat}};
    for (const auto &c : cp)
        if (e == hipSuccess) e = hipMemcpyAsync(c.dst->ptr, c.src->ptr, c.src->bytes, hipMemcpyDeviceToDevice, st);
    // every slot starts free
    std::vector<int> idle(cfg->capacity, SMR_RAFF_IDLE);
    if (e == hipSuccess) e = hipMemcpyAsync(r->status.ptr, idle.data(), sizeof(int) * idle.size(), hipMemcpyHostToDevice, st);
    std::vector<float4> p0(n);
    if (e == hipSuccess) e = hipMemcpyAsync(p0.data(), tmpl->pos.ptr, sizeof(float4) * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipStream_t own = nullptr;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking);
    if (e != hipSuccess) return fail_r(e);
    r->bk_host = tmpl->bk_host;
    r->neigh_host = tmpl->neigh_host;
    r->have_neighs = r->have_bk = r->have_state = true;
    auto *s = new smr_raff_server();
    s->r = r;
    s->stream = own;
    s->cfg = *cfg;
    s->inv_mass.resize(n);
    for (int i = 0; i < n; i++) s->inv_mass[i] = p0[i].w;
    s->slot_id.assign(cfg->capacity, -1);
    for (int k = cfg->capacity - 1; k >= 0; k--) s->free_slots.push_back(k);
    s->hsteps.assign(cfg->capacity, 0);
    s->hstatus.assign(cfg->capacity, SMR_RAFF_IDLE);
    s->hmetric.assign(cfg->capacity, 0.f);
    s->t0 = clk::now();
    s->worker = std::thread(raff_worker, s);
    *out = s;
    return SMR_OK;
}

extern "C" int smr_raff_server_submit(smr_raff_server *s, const float *pos3, const float *quat, long long *id) {
    if (!s || !pos3 || !id) return fail(SMR_ERR_INVALID, "smr_raff_server_submit: NULL argument");
    const int n = s->r->n;
    auto *q = new RaffReq();
    q->pos4.resize(n);
    q->quat.resize(n);
    const float nan = std::nanf("");
    for (int i = 0; i < n; i++) {
        const float w = s->inv_mass[i];
        if (w <= 1e-12f) {
            q->pos4[i] = make_float4(nan, nan, nan, w);
        } else {
            if (!std::isfinite(pos3[3 * i]) || !std::isfinite(pos3[3 * i + 1]) || !std::isfinite(pos3[3 * i + 2])) {
                delete q;
                return fail(SMR_ERR_INVALID, "smr_raff_server_submit: position of atom " + std::to_string(i) +
                                                 " is not finite");
            }
            q->pos4[i] = make_float4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], w);
        }
        q->quat[i] = quat ? make_float4(quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3])
                          : make_float4(0.f, 0.f, 0.f, 1.f);
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop)
            return delete q, fail(SMR_ERR_STATE, "smr_raff_server_submit: the server stopped (" + s->worker_error + ")");
        q->id = s->next_id++;
        q->t_submit = now_s(s);
        s->reqs[q->id] = q;
        s->queue.push_back(q);
        s->submitted++;
        *id = q->id;
    }
    s->cv.notify_one();
    return SMR_OK;
}

extern "C" int smr_raff_server_poll(smr_raff_server *s, long long id, int *done, int *steps, int *status, float *metric,
                                    float *pos3, double *t_submit, double *t_done) {
    if (!s || !done) return fail(SMR_ERR_INVALID, "smr_raff_server_poll: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(SMR_ERR_INVALID, "smr_raff_server_poll: unknown request id " + std::to_string(id));
    RaffReq *q = it->second;
    *done = q->done ? 1 : 0;
    if (!q->done) return SMR_OK;
    if (steps) *steps = q->steps;
    if (status) *status = q->status;
    if (metric) *metric = q->metric;
    if (pos3) memcpy(pos3, q->pos_out.data(), sizeof(float) * q->pos_out.size());
    if (t_submit) *t_submit = q->t_submit;
    if (t_done) *t_done = q->t_done;
    s->reqs.erase(it);
    delete q;
    return SMR_OK;
}

extern "C" int smr_raff_server_drain(smr_raff_server *s) {
    if (!s) return fail(SMR_ERR_INVALID, "smr_raff_server_drain: NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->stop || (s->queue.empty() && s->resident == 0); });
    if (!s->worker_error.empty()) return fail(SMR_ERR_HIP, s->worker_error);
    return SMR_OK;
}

extern "C" int smr_raff_server_stats(smr_raff_server *s, double *out) {
    if (!s || !out) return fail(SMR_ERR_INVALID, "smr_raff_server_stats: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = s->submitted;
    out[1] = s->retired;
    out[2] = s->resident;
    out[3] = (double)s->queue.size();
    out[4] = s->chunks;
    out[5] = s->peak;
    out[6] = s->device_s;
    return SMR_OK;
}

extern "C" int smr_raff_server_destroy(smr_raff_server *s) {
    if (!s) return SMR_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto &kv : s->reqs) delete kv.second;
    (void)hipSetDevice(s->r->ctx->device);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    smr_raff_destroy(s->r);
    delete s;
    return SMR_OK;
}
