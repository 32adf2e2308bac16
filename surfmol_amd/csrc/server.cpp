// server.cpp -- continuous-batching relax server (smr_server_* of
// include/surfmol_relax.h).
//
// Device side: one set of arenas sized at creation from a fraction of the
// free HBM -- per-atom pools (state, topology tables, FP64 parameters),
// per-term pools (angles, dihedrals, inversions, gather lists) -- and a slot
// table (TopoDesc / SysDesc / SysState per slot).  A request occupies one
// slot and one range in every arena; ranges come from first-fit free lists,
// so requests of any size are admitted while their memory is free.
//
// Host side: submit() validates and packs a request's tables in the
// caller's thread and queues it.  A worker thread loops:
//   admit queued requests into free slots / arena ranges (FIFO);
//   upload their tables, positions and FIRE state with one staged copy;
//   run one chunk: a few FIRE steps of every resident request (chunk_steps
//     for the smallest size class, fewer for the larger ones), one
//     persistent launch per size class (launch_plans);
//   read the FIRE states back; for requests that converged or hit the step
//     cap, evaluate FP64 energies, copy their positions out, free their
//     slot and ranges, and mark them done.
// Requests that arrive while a chunk runs join at the next chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "../../include/surfmol_relax.h"
#include "smr_internal.h"

using namespace smr;

namespace {

using clk = std::chrono::steady_clock;

// first-fit range allocator with coalescing free list
class RangeAlloc {
  public:
    void init(size_t cap) {
        free_.clear();
        if (cap) free_[0] = cap;
        cap_ = cap;
        used_ = 0;
    }
    bool alloc(size_t n, size_t &off) {
        off = 0;
        if (n == 0) return true;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const size_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            used_ += n;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        if (n == 0) return;
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t cap() const { return cap_; }
    size_t used() const { return used_; }

  private:
    std::map<size_t, size_t> free_;
    size_t cap_ = 0, used_ = 0;
};

struct Request {
    long long id = 0;
    HostPools hp;            // this request's packed tables (offsets from 0)
    TopoDesc td;             // counts; arena offsets filled at admission
    std::vector<float4> pos0;
    double org[3] = {0, 0, 0};
    int iux = 0, iuy = 0;
    int slot = -1;
    size_t a_off = 0, an_off = 0, d_off = 0, i_off = 0, f_off = 0;
    double t_submit = 0, t_done = 0;
    bool done = false;
    smr_relax_result res;
    std::vector<double> out;
};

// per-atom / per-term arena sizes relative to the atom capacity (upper
// bounds of fused-ring and chain molecules: <= 3 angles, 6 dihedrals,
// 3 inversions per atom)
constexpr size_t kAngPerAtom = 3, kDihPerAtom = 6, kInvPerAtom = 3;
constexpr size_t kPiecesPerAtom = 4 * kDihPerAtom + 4 * kInvPerAtom + 3 * kAngPerAtom;
constexpr size_t kBytesPerAtom = 3 * 16 + 5 * 16 + 8 + 4 * 32 + 256 + 4;   // state, tables, FP64 tables, stencil
constexpr size_t kBytesPerAng = 4 + 8 + 16 + 32, kBytesPerDih = 8 + 8 + 16 + 32, kBytesPerInv = 8 + 8 + 16 + 32;
constexpr size_t kUnitBytes = kBytesPerAtom + kAngPerAtom * kBytesPerAng + kDihPerAtom * kBytesPerDih +
                              kInvPerAtom * kBytesPerInv + kPiecesPerAtom * 2;

}  // namespace

struct smr_server {
    smr_ctx *ctx = nullptr;
    const smr_gridff *grid = nullptr;
    smr_server_config cfg;
    StreamSet ss;
    LaunchSet ls;
    DevTables tab;
    FireConst fc;
    DevBuf pos, vel, force, stencil, skey, ng, bk, bl, req, plq, a2f, bk64, bl64, req64, plq64;
    DevBuf ang_ng, ang_par, ang64, dih_ng, dih_par, dih64, inv_ng, inv_par, inv64, a2f_idx;
    DevBuf ang_dst, dih_dst, inv_dst;
    DevBuf topo, sys, state, work, list, dstage, dout;
    RangeAlloc ra_atoms, ra_ang, ra_dih, ra_inv, ra_pc;
    size_t arena_bytes = 0;
    void *hstage = nullptr;
    size_t hstage_bytes = 0;
    void *hout = nullptr;
    size_t hout_bytes = 0;
    SysState *hstate = nullptr;
    int *hwork = nullptr;
    std::vector<int> free_slots;
    std::vector<Request *> slot_req;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<long long, Request *> reqs;
    std::thread worker;
    bool stop = false;
    long long next_id = 1;
    long long n_submitted = 0, n_retired = 0, chunks = 0;
    int n_resident = 0, peak = 0;
    double dev_s = 0.0;
    int err_code = 0;
    std::string err_msg;
    clk::time_point t0;
};

namespace {

double now_s(const smr_server *S) { return std::chrono::duration<double>(clk::now() - S->t0).count(); }

hipError_t grow_pinned(void *&p, size_t &have, size_t need) {
    if (have >= need) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    have = 0;
    size_t n = std::max<size_t>(need, 1 << 20);
    n = std::max(n, need + need / 2);
    hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
    if (e == hipSuccess) have = n;
    return e;
}

hipError_t grow_dev(DevBuf &b, size_t need) {
    if (b.bytes >= need) return hipSuccess;
    return b.alloc(std::max(need, b.bytes + b.bytes / 2));
}

// reserve arena ranges for r (all or nothing)
bool place(smr_server *S, Request *r) {
    const TopoDesc &t = r->td;
    size_t a, an, d, i, f;
    if (!S->ra_atoms.alloc(t.natoms, a)) return false;
    if (!S->ra_ang.alloc(t.nangles, an)) { S->ra_atoms.release(a, t.natoms); return false; }
    if (!S->ra_dih.alloc(t.ndih, d)) {
        S->ra_atoms.release(a, t.natoms); S->ra_ang.release(an, t.nangles);
        return false;
    }
    if (!S->ra_inv.alloc(t.ninv, i)) {
        S->ra_atoms.release(a, t.natoms); S->ra_ang.release(an, t.nangles); S->ra_dih.release(d, t.ndih);
        return false;
    }
    if (!S->ra_pc.alloc(t.nf, f)) {
        S->ra_atoms.release(a, t.natoms); S->ra_ang.release(an, t.nangles); S->ra_dih.release(d, t.ndih);
        S->ra_inv.release(i, t.ninv);
        return false;
    }
    r->a_off = a; r->an_off = an; r->d_off = d; r->i_off = i; r->f_off = f;
    return true;
}

void unplace(smr_server *S, Request *r) {
    const TopoDesc &t = r->td;
    S->ra_atoms.release(r->a_off, t.natoms);
    S->ra_ang.release(r->an_off, t.nangles);
    S->ra_dih.release(r->d_off, t.ndih);
    S->ra_inv.release(r->i_off, t.ninv);
    S->ra_pc.release(r->f_off, t.nf);
}

// staging of many small host pieces into one pinned buffer + descriptors
struct Stager {
    std::vector<unsigned char> bytes;
    std::vector<CopyDesc> desc;
    std::vector<size_t> src_off;    // host offsets, turned into device addresses at flush
    void add(const void *src, size_t n, void *dst) {
        if (n == 0) return;
        size_t off = (bytes.size() + 15) & ~(size_t)15;
        bytes.resize(off + n);
        memcpy(bytes.data() + off, src, n);
        CopyDesc c;
        c.src = off;
        c.dst = (unsigned long long)dst;
        c.bytes = n;
        desc.push_back(c);
    }
};

int flush(smr_server *S, Stager &g) {
    if (g.desc.empty()) return SMR_OK;
    const size_t dbytes = g.desc.size() * sizeof(CopyDesc);
    const size_t data_off = (dbytes + 255) & ~(size_t)255;
    const size_t total = data_off + g.bytes.size();
    hipError_t e = hipStreamSynchronize(S->ss.main);   // the pinned buffer may still feed a copy
    if (e == hipSuccess) e = grow_pinned(S->hstage, S->hstage_bytes, total);
    if (e == hipSuccess) e = grow_dev(S->dstage, total);
    if (e != hipSuccess) return hip_fail(e, "server staging buffers");
    const unsigned long long dbase = (unsigned long long)S->dstage.ptr + data_off;
    for (auto &c : g.desc) c.src += dbase;
    memcpy(S->hstage, g.desc.data(), dbytes);
    memcpy((unsigned char *)S->hstage + data_off, g.bytes.data(), g.bytes.size());
    e = hipMemcpyAsync(S->dstage.ptr, S->hstage, total, hipMemcpyHostToDevice, S->ss.main);
    if (e == hipSuccess) e = launch_copy_desc((const CopyDesc *)S->dstage.ptr, (int)g.desc.size(), S->ss.main);
    if (e != hipSuccess) return hip_fail(e, "server upload");
    return SMR_OK;
}

template <class T>
T *at(DevBuf &b, size_t i) {
    return (T *)b.ptr + i;
}

int upload(smr_server *S, const std::vector<Request *> &admit) {
    Stager g;
    std::vector<float4> zeros;
    for (Request *r : admit) {
        HostPools &h = r->hp;
        TopoDesc td = r->td;
        const int N = td.natoms;
        td.atom_off = (int)r->a_off;
        td.ang_off = (int)r->an_off;
        td.dih_off = (int)r->d_off;
        td.inv_off = (int)r->i_off;
        td.a2f_off = (int)r->f_off;
        SysDesc sd;
        sd.topo = r->slot;
        sd.atom_base = (int)r->a_off;
        sd.iux = r->iux;
        sd.iuy = r->iuy;
        sd.gx = sd.gy = sd.gz = sd.pad = 0;
        SysState st;
        memset(&st, 0, sizeof(st));
        st.dt = S->cfg.fire.dt0;
        st.alpha = S->cfg.fire.alpha0;
        st.status = SMR_SYS_RUNNING;
        const size_t a = r->a_off;
        g.add(h.ng.data(), N * sizeof(int4), at<int4>(S->ng, a));
        g.add(h.bk.data(), N * sizeof(float4), at<float4>(S->bk, a));
        g.add(h.bl.data(), N * sizeof(float4), at<float4>(S->bl, a));
        g.add(h.req.data(), N * sizeof(float4), at<float4>(S->req, a));
        g.add(h.plq.data(), N * sizeof(float4), at<float4>(S->plq, a));
        g.add(h.a2f.data(), N * sizeof(int2), at<int2>(S->a2f, a));
        g.add(h.bk64.data(), N * sizeof(double4), at<double4>(S->bk64, a));
        g.add(h.bl64.data(), N * sizeof(double4), at<double4>(S->bl64, a));
        g.add(h.req64.data(), N * sizeof(double4), at<double4>(S->req64, a));
        g.add(h.plq64.data(), N * sizeof(double4), at<double4>(S->plq64, a));
        g.add(r->pos0.data(), N * sizeof(float4), at<float4>(S->pos, a));
        zeros.assign(N, make_float4(0.f, 0.f, 0.f, 0.f));
        g.add(zeros.data(), N * sizeof(float4), at<float4>(S->vel, a));
        // the slot's stored surface stencils belong to the previous request
        std::vector<int> nokey(N, -1);
        g.add(nokey.data(), N * sizeof(int), at<int>(S->skey, a));
        g.add(h.ang_ng.data(), td.nangles * sizeof(unsigned), at<unsigned>(S->ang_ng, r->an_off));
        g.add(h.ang_par.data(), td.nangles * sizeof(float4), at<float4>(S->ang_par, r->an_off));
        g.add(h.ang64.data(), td.nangles * sizeof(double4), at<double4>(S->ang64, r->an_off));
        g.add(h.dih_ng.data(), td.ndih * sizeof(uint2), at<uint2>(S->dih_ng, r->d_off));
        g.add(h.dih_par.data(), td.ndih * sizeof(float4), at<float4>(S->dih_par, r->d_off));
        g.add(h.dih64.data(), td.ndih * sizeof(double4), at<double4>(S->dih64, r->d_off));
        g.add(h.inv_ng.data(), td.ninv * sizeof(uint2), at<uint2>(S->inv_ng, r->i_off));
        g.add(h.inv_par.data(), td.ninv * sizeof(float4), at<float4>(S->inv_par, r->i_off));
        g.add(h.inv64.data(), td.ninv * sizeof(double4), at<double4>(S->inv64, r->i_off));
        g.add(h.a2f_idx.data(), td.nf * sizeof(unsigned short), at<unsigned short>(S->a2f_idx, r->f_off));
        g.add(h.ang_dst.data(), td.nangles * sizeof(uint2), at<uint2>(S->ang_dst, r->an_off));
        g.add(h.dih_dst.data(), td.ndih * sizeof(uint2), at<uint2>(S->dih_dst, r->d_off));
        g.add(h.inv_dst.data(), td.ninv * sizeof(uint2), at<uint2>(S->inv_dst, r->i_off));
        g.add(&td, sizeof(TopoDesc), at<TopoDesc>(S->topo, r->slot));
        g.add(&sd, sizeof(SysDesc), at<SysDesc>(S->sys, r->slot));
        g.add(&st, sizeof(SysState), at<SysState>(S->state, r->slot));
        r->td = td;
    }
    return flush(S, g);
}

// one chunk of every resident request, then retirement
int run_chunk(smr_server *S) {
    std::vector<SysSize> sizes;
    for (size_t k = 0; k < S->slot_req.size(); k++) {
        const Request *r = S->slot_req[k];
        if (!r) continue;
        sizes.push_back(SysSize{(int)k, r->td.natoms, r->td.nf, r->td.nangles, r->td.ndih, r->td.ninv});
    }
    if (sizes.empty()) return SMR_OK;
    std::vector<ClassPlan> plans;
    std::vector<int> work;
    build_plans(sizes, plans, work);
    // a chunk ends when its slowest launch ends, and a step of a larger
    // class costs more (its pair sweep grows with the atom count): class c
    // takes chunk_steps * 2 / (2 + c) steps (64, 42, 32, 25, 21 of 64), so
    // the small requests, most of the traffic, wait less for the large ones'
    // steps. Halving per class instead starved the large requests (p99 up
    // by a third); chunk boundaries change no result
    for (auto &p : plans) {
        p.shape = class_shape(S->ctx, p, S->grid != nullptr, 0, -1, -1);
        p.steps = std::max(1, S->cfg.chunk_steps * 2 / (2 + p.cls));
    }
    hipError_t e = hipStreamSynchronize(S->ss.main);   // hwork may still feed the previous copy
    if (e != hipSuccess) return hip_fail(e, "server chunk");
    memcpy(S->hwork, work.data(), work.size() * sizeof(int));
    e = hipMemcpyAsync(S->work.ptr, S->hwork, work.size() * sizeof(int), hipMemcpyHostToDevice, S->ss.main);
    if (e != hipSuccess) return hip_fail(e, "server work list");
    KernelArgs A;
    memset(&A, 0, sizeof(A));
    A.tab = S->tab;
    A.fire = S->fc;
    A.pos = (float4 *)S->pos.ptr;
    A.vel = (float4 *)S->vel.ptr;
    A.force = (float4 *)S->force.ptr;
    A.state = (SysState *)S->state.ptr;
    A.work_list = (const int *)S->work.ptr;
    A.stencil = (float4 *)S->stencil.ptr;
    A.stencil_key = (int *)S->skey.ptr;
    A.steps_per_launch = S->cfg.chunk_steps;
    A.mode = KMODE_FIRE;
    float ms = 0.f;
    int rc = launch_plans(S->ss, S->ls, plans, A, &ms);
    if (rc) return rc;
    S->chunks++;
    S->dev_s += ms * 1e-3;
    const size_t nslot = S->slot_req.size();
    e = hipMemcpyAsync(S->hstate, S->state.ptr, nslot * sizeof(SysState), hipMemcpyDeviceToHost, S->ss.main);
    if (e == hipSuccess) e = hipStreamSynchronize(S->ss.main);
    if (e != hipSuccess) return hip_fail(e, "server state readback");
    std::vector<int> fin;
    for (size_t k = 0; k < nslot; k++)
        if (S->slot_req[k] && S->hstate[k].status != SMR_SYS_RUNNING) fin.push_back((int)k);
    if (fin.empty()) return SMR_OK;
    // FP64 energies and final positions of the retiring requests
    memcpy(S->hwork, fin.data(), fin.size() * sizeof(int));
    e = hipMemcpyAsync(S->list.ptr, S->hwork, fin.size() * sizeof(int), hipMemcpyHostToDevice, S->ss.main);
    if (e == hipSuccess)
        e = launch_energy64(S->tab, (const float4 *)S->pos.ptr, (SysState *)S->state.ptr, (const int *)S->list.ptr,
                            (int)fin.size(), S->ss.main);
    if (e != hipSuccess) return hip_fail(e, "server energies");
    size_t nat = 0;
    for (int k : fin) nat += S->slot_req[k]->td.natoms;
    const size_t out_bytes = nat * sizeof(float4) + fin.size() * sizeof(SysState);
    e = grow_dev(S->dout, out_bytes);
    if (e == hipSuccess) e = grow_pinned(S->hout, S->hout_bytes, out_bytes);
    if (e != hipSuccess) return hip_fail(e, "server output buffers");
    // descriptors: positions then states, packed back to back in dout
    std::vector<CopyDesc> desc;
    size_t off = 0;
    for (int k : fin) {
        const Request *r = S->slot_req[k];
        desc.push_back(CopyDesc{(unsigned long long)at<float4>(S->pos, r->a_off),
                                (unsigned long long)((unsigned char *)S->dout.ptr + off),
                                (unsigned long long)(r->td.natoms * sizeof(float4))});
        off += r->td.natoms * sizeof(float4);
    }
    for (int k : fin) {
        desc.push_back(CopyDesc{(unsigned long long)at<SysState>(S->state, k),
                                (unsigned long long)((unsigned char *)S->dout.ptr + off), sizeof(SysState)});
        off += sizeof(SysState);
    }
    // the descriptor list travels through the staging buffer
    const size_t dbytes = desc.size() * sizeof(CopyDesc);
    e = grow_dev(S->dstage, dbytes);
    if (e == hipSuccess) e = grow_pinned(S->hstage, S->hstage_bytes, dbytes);
    if (e != hipSuccess) return hip_fail(e, "server staging buffers");
    memcpy(S->hstage, desc.data(), dbytes);
    e = hipMemcpyAsync(S->dstage.ptr, S->hstage, dbytes, hipMemcpyHostToDevice, S->ss.main);
    if (e == hipSuccess) e = launch_copy_desc((const CopyDesc *)S->dstage.ptr, (int)desc.size(), S->ss.main);
    if (e == hipSuccess) e = hipMemcpyAsync(S->hout, S->dout.ptr, out_bytes, hipMemcpyDeviceToHost, S->ss.main);
    if (e == hipSuccess) e = hipStreamSynchronize(S->ss.main);
    if (e != hipSuccess) return hip_fail(e, "server retirement copies");
    const double t = now_s(S);
    const float4 *hp = (const float4 *)S->hout;
    const SysState *hs = (const SysState *)((const unsigned char *)S->hout + nat * sizeof(float4));
    std::vector<Request *> done;
    for (size_t q = 0; q < fin.size(); q++) {
        Request *r = S->slot_req[fin[q]];
        const SysState &st = hs[q];
        r->res.steps = st.steps;
        r->res.status = st.status;
        r->res.fmax = st.fmax;
        double E = 0.0;
        for (int c = 0; c < 6; c++) { r->res.terms[c] = st.E[c]; E += st.E[c]; }
        r->res.energy = E;
        const int N = r->td.natoms;
        r->out.resize(3 * (size_t)N);
        for (int i = 0; i < N; i++) {
            r->out[3 * i] = (double)hp[i].x + r->org[0];
            r->out[3 * i + 1] = (double)hp[i].y + r->org[1];
            r->out[3 * i + 2] = (double)hp[i].z + r->org[2];
        }
        hp += N;
        r->t_done = t;
        unplace(S, r);
        done.push_back(r);
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (Request *r : done) {
            S->slot_req[r->slot] = nullptr;
            S->free_slots.push_back(r->slot);
            r->slot = -1;
            r->done = true;
            r->hp = HostPools();   // tables are on the device no longer needed on the host
            S->n_resident--;
            S->n_retired++;
        }
    }
    S->cv_done.notify_all();
    return SMR_OK;
}

void worker_main(smr_server *S) {
    (void)hipSetDevice(S->ctx->device);
    for (;;) {
        std::vector<Request *> admit;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty() || S->n_resident > 0; });
            if (S->stop) break;
            while (!S->pending.empty() && !S->free_slots.empty()) {
                Request *r = S->pending.front();
                if (!place(S, r)) break;   // arenas full: the head waits for retirements
                S->pending.pop_front();
                r->slot = S->free_slots.back();
                S->free_slots.pop_back();
                S->slot_req[r->slot] = r;
                admit.push_back(r);
                S->n_resident++;
            }
            S->peak = std::max(S->peak, S->n_resident);
        }
        int rc = admit.empty() ? SMR_OK : upload(S, admit);
        if (rc == SMR_OK) rc = run_chunk(S);
        if (rc != SMR_OK) {
            std::lock_guard<std::mutex> lk(S->mu);
            S->err_code = rc;
            S->err_msg = "serving loop stopped: " + g_err;
            S->stop = true;
            S->cv_done.notify_all();
            break;
        }
    }
}

void free_host(smr_server *S) {
    if (S->hstage) (void)hipHostFree(S->hstage);
    if (S->hout) (void)hipHostFree(S->hout);
    if (S->hstate) (void)hipHostFree(S->hstate);
    if (S->hwork) (void)hipHostFree(S->hwork);
    S->hstage = S->hout = nullptr;
    S->hstate = nullptr;
    S->hwork = nullptr;
}

int server_error(smr_server *S) {
    if (!S->err_code) return SMR_OK;
    return fail(S->err_code, S->err_msg);
}

}  // namespace

extern "C" int smr_server_create(smr_ctx *c, const smr_gridff *grid, const smr_server_config *cfg, smr_server **out) {
    if (!c || !cfg || !out) return fail(SMR_ERR_INVALID, "smr_server_create: NULL argument");
    *out = nullptr;
    if (cfg->max_resident <= 0 || cfg->max_resident > (1 << 20))
        return fail(SMR_ERR_INVALID, "smr_server_create: max_resident must be in [1, 2^20]");
    if (cfg->max_atoms <= 0 || cfg->max_atoms > SMR_MAX_ATOMS)
        return fail(SMR_ERR_INVALID, "smr_server_create: max_atoms must be in [1, " + std::to_string(SMR_MAX_ATOMS) + "]");
    if (!(cfg->hbm_fraction > 0.0 && cfg->hbm_fraction <= 1.0))
        return fail(SMR_ERR_INVALID, "smr_server_create: hbm_fraction must be in (0, 1]");
    if (cfg->chunk_steps <= 0 || cfg->max_steps <= 0 || !(cfg->fconv > 0.0))
        return fail(SMR_ERR_INVALID, "smr_server_create: need chunk_steps > 0, max_steps > 0, fconv > 0");
    int rc = check_fire(&cfg->fire, "smr_server_create");
    if (rc) return rc;
    if (cfg->ff.surface && !grid) return fail(SMR_ERR_INVALID, "smr_server_create: surface requested but grid == NULL");
    if (cfg->ff.inter_cutoff != 0.0 || cfg->ff.pbc_lx != 0.0 || cfg->ff.pbc_ly != 0.0)
        return fail(SMR_ERR_UNSUPPORTED, "smr_server_create: requests are independent molecules (inter_cutoff must be 0)");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    size_t free_b = 0, total_b = 0;
    e = hipMemGetInfo(&free_b, &total_b);
    if (e != hipSuccess) return hip_fail(e, "hipMemGetInfo");
    const double budget = cfg->hbm_fraction * (double)free_b;
    const size_t want = (size_t)cfg->max_resident * (size_t)cfg->max_atoms;
    // arena offsets are 32-bit in the device tables
    const size_t cap = std::min({want, (size_t)(budget / (double)kUnitBytes), (size_t)0x7fffffff / kPiecesPerAtom});
    if (cap < (size_t)cfg->max_atoms)
        return fail(SMR_ERR_NOMEM, "smr_server_create: hbm_fraction leaves room for " + std::to_string(cap) +
                                       " atoms, fewer than max_atoms");
    auto *S = new smr_server();
    S->ctx = c;
    ctx_retain(c);
    S->grid = cfg->ff.surface ? grid : nullptr;
    S->cfg = *cfg;
    S->t0 = clk::now();
    S->ls.use_graph = false;   // the work list changes every chunk
    const size_t nr = (size_t)cfg->max_resident;
    const size_t ca = cap, cn = cap * kAngPerAtom, cd = cap * kDihPerAtom, ci = cap * kInvPerAtom,
                 cf = cap * kPiecesPerAtom;
    auto fail_create = [&](hipError_t err, const char *what) {
        S->ss.destroy();
        free_host(S);
        delete S;
        ctx_release(c);
        return hip_fail(err, what);
    };
    e = S->ss.create();
    if (e != hipSuccess) return fail_create(e, "smr_server_create: streams");
#define ALLOC(buf, bytes) \
    if ((e = S->buf.alloc(bytes)) != hipSuccess) return fail_create(e, "smr_server_create: arena " #buf);
    ALLOC(pos, ca * sizeof(float4)) ALLOC(vel, ca * sizeof(float4)) ALLOC(force, ca * sizeof(float4))
    ALLOC(stencil, ca * 16 * sizeof(float4)) ALLOC(skey, ca * sizeof(int))
    ALLOC(ng, ca * sizeof(int4)) ALLOC(bk, ca * sizeof(float4)) ALLOC(bl, ca * sizeof(float4))
    ALLOC(req, ca * sizeof(float4)) ALLOC(plq, ca * sizeof(float4)) ALLOC(a2f, ca * sizeof(int2))
    ALLOC(bk64, ca * sizeof(double4)) ALLOC(bl64, ca * sizeof(double4)) ALLOC(req64, ca * sizeof(double4))
    ALLOC(plq64, ca * sizeof(double4))
    ALLOC(ang_ng, cn * sizeof(unsigned)) ALLOC(ang_par, cn * sizeof(float4)) ALLOC(ang64, cn * sizeof(double4))
    ALLOC(dih_ng, cd * sizeof(uint2)) ALLOC(dih_par, cd * sizeof(float4)) ALLOC(dih64, cd * sizeof(double4))
    ALLOC(inv_ng, ci * sizeof(uint2)) ALLOC(inv_par, ci * sizeof(float4)) ALLOC(inv64, ci * sizeof(double4))
    ALLOC(a2f_idx, cf * sizeof(unsigned short))
    ALLOC(ang_dst, cn * sizeof(uint2)) ALLOC(dih_dst, cd * sizeof(uint2)) ALLOC(inv_dst, ci * sizeof(uint2))
    ALLOC(topo, nr * sizeof(TopoDesc)) ALLOC(sys, nr * sizeof(SysDesc)) ALLOC(state, nr * sizeof(SysState))
    ALLOC(work, nr * sizeof(int)) ALLOC(list, nr * sizeof(int))
#undef ALLOC
    S->arena_bytes = cap * kUnitBytes;
    void *hp = nullptr;
    if ((e = hipHostMalloc(&hp, nr * sizeof(SysState), hipHostMallocDefault)) != hipSuccess)
        return fail_create(e, "smr_server_create: pinned state buffer");
    S->hstate = (SysState *)hp;
    if ((e = hipHostMalloc(&hp, nr * sizeof(int), hipHostMallocDefault)) != hipSuccess)
        return fail_create(e, "smr_server_create: pinned work buffer");
    S->hwork = (int *)hp;
    S->ra_atoms.init(ca);
    S->ra_ang.init(cn);
    S->ra_dih.init(cd);
    S->ra_inv.init(ci);
    S->ra_pc.init(cf);
    S->slot_req.assign(nr, nullptr);
    for (int k = (int)nr - 1; k >= 0; k--) S->free_slots.push_back(k);
    // device tables over the arenas
    DevTables &T = S->tab;
    memset(&T, 0, sizeof(T));
    T.topo = (const TopoDesc *)S->topo.ptr;
    T.sys = (const SysDesc *)S->sys.ptr;
    T.ng = (const int4 *)S->ng.ptr;
    T.bk = (const float4 *)S->bk.ptr;
    T.bl = (const float4 *)S->bl.ptr;
    T.req = (const float4 *)S->req.ptr;
    T.plq = (const float4 *)S->plq.ptr;
    T.a2f = (const int2 *)S->a2f.ptr;
    T.ang_ng = (const unsigned *)S->ang_ng.ptr;
    T.ang_par = (const float4 *)S->ang_par.ptr;
    T.dih_ng = (const uint2 *)S->dih_ng.ptr;
    T.dih_par = (const float4 *)S->dih_par.ptr;
    T.inv_ng = (const uint2 *)S->inv_ng.ptr;
    T.inv_par = (const float4 *)S->inv_par.ptr;
    T.a2f_idx = (const unsigned short *)S->a2f_idx.ptr;
    T.ang_dst = (const uint2 *)S->ang_dst.ptr;
    T.dih_dst = (const uint2 *)S->dih_dst.ptr;
    T.inv_dst = (const uint2 *)S->inv_dst.ptr;
    T.bk64 = (const double4 *)S->bk64.ptr;
    T.bl64 = (const double4 *)S->bl64.ptr;
    T.req64 = (const double4 *)S->req64.ptr;
    T.plq64 = (const double4 *)S->plq64.ptr;
    T.ang64 = (const double4 *)S->ang64.ptr;
    T.dih64 = (const double4 *)S->dih64.ptr;
    T.inv64 = (const double4 *)S->inv64.ptr;
    if (S->grid) {
        const smr_grid_desc &d = S->grid->desc;
        T.grid = (const float4 *)S->grid->coef.ptr;
        T.gnx = d.nx; T.gny = d.ny; T.gnz = d.nz;
        T.ginv_x = (float)(1.0 / d.step[0]);
        T.ginv_y = (float)(1.0 / d.step[1]);
        T.ginv_z = (float)(1.0 / d.step[2]);
        T.ginv64_x = 1.0 / d.step[0];
        T.ginv64_y = 1.0 / d.step[1];
        T.ginv64_z = 1.0 / d.step[2];
    } else {
        T.grid = nullptr;
        T.gnx = T.gny = T.gnz = 1;
        T.ginv_x = T.ginv_y = T.ginv_z = 1.f;
        T.ginv64_x = T.ginv64_y = T.ginv64_z = 1.0;
    }
    T.r2damp = (float)(cfg->ff.rdamp * cfg->ff.rdamp);
    T.r2damp64 = cfg->ff.rdamp * cfg->ff.rdamp;
    T.use_pairs = cfg->ff.nonbonded ? 1 : 0;
    S->fc = make_fire_const(&cfg->fire, cfg->max_steps, cfg->fconv);
    S->worker = std::thread(worker_main, S);
    *out = S;
    return SMR_OK;
}

extern "C" int smr_server_submit(smr_server *S, const smr_topology *topo, const double *apos, long long *id) {
    if (!S || !topo || !apos || !id) return fail(SMR_ERR_INVALID, "smr_server_submit: NULL argument");
    auto *r = new Request();
    const double alpha = S->grid ? S->grid->desc.alpha_morse : 0.0;
    int rc = pack_topology(*topo, 0, alpha, S->grid != nullptr, r->hp, r->td);
    if (rc) { delete r; return rc; }
    const TopoDesc &t = r->td;
    if (t.natoms > S->cfg.max_atoms) {
        delete r;
        return fail(SMR_ERR_UNSUPPORTED, "smr_server_submit: " + std::to_string(t.natoms) + " atoms > max_atoms " +
                                             std::to_string(S->cfg.max_atoms));
    }
    if ((size_t)t.natoms > S->ra_atoms.cap() || (size_t)t.nangles > S->ra_ang.cap() || (size_t)t.ndih > S->ra_dih.cap() ||
        (size_t)t.ninv > S->ra_inv.cap() || (size_t)t.nf > S->ra_pc.cap()) {
        delete r;
        return fail(SMR_ERR_NOMEM, "smr_server_submit: the request's tables exceed the server's arenas");
    }
    const int N = t.natoms;
    int g[3];
    system_origin(apos, N, S->grid, S->grid != nullptr, r->org, r->iux, r->iuy, g);
    r->pos0.resize(N);
    for (int i = 0; i < N; i++) {
        const double x = apos[3 * i] - r->org[0], y = apos[3 * i + 1] - r->org[1], z = apos[3 * i + 2] - r->org[2];
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z)) {
            delete r;
            return fail(SMR_ERR_INVALID, "smr_server_submit: atom " + std::to_string(i) + " is not finite");
        }
        r->pos0[i] = make_float4((float)x, (float)y, (float)z, 0.f);
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->err_code) { delete r; return server_error(S); }
        if (S->stop) { delete r; return fail(SMR_ERR_STATE, "smr_server_submit: server is stopping"); }
        r->id = S->next_id++;
        r->t_submit = now_s(S);
        S->reqs[r->id] = r;
        S->pending.push_back(r);
        S->n_submitted++;
        *id = r->id;
    }
    S->cv_work.notify_one();
    return SMR_OK;
}

extern "C" int smr_server_poll(smr_server *S, long long id, int *done, smr_relax_result *result, double *apos_out,
                               double *t_submit, double *t_done) {
    if (!S || !done) return fail(SMR_ERR_INVALID, "smr_server_poll: NULL argument");
    Request *r = nullptr;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->err_code) return server_error(S);
        auto it = S->reqs.find(id);
        if (it == S->reqs.end()) return fail(SMR_ERR_INVALID, "smr_server_poll: unknown request id " + std::to_string(id));
        r = it->second;
        if (!r->done) {
            *done = 0;
            return SMR_OK;
        }
        S->reqs.erase(it);
    }
    *done = 1;
    if (result) *result = r->res;
    if (apos_out) std::copy(r->out.begin(), r->out.end(), apos_out);
    if (t_submit) *t_submit = r->t_submit;
    if (t_done) *t_done = r->t_done;
    delete r;
    return SMR_OK;
}

extern "C" int smr_server_drain(smr_server *S) {
    if (!S) return fail(SMR_ERR_INVALID, "smr_server_drain: NULL argument");
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->err_code != 0 || (S->pending.empty() && S->n_resident == 0); });
    return server_error(S);
}

extern "C" int smr_server_stats(smr_server *S, double *out) {
    if (!S || !out) return fail(SMR_ERR_INVALID, "smr_server_stats: NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    out[0] = (double)S->n_submitted;
    out[1] = (double)S->n_retired;
    out[2] = (double)S->n_resident;
    out[3] = (double)S->pending.size();
    out[4] = (double)S->chunks;
    out[5] = (double)S->peak;
    out[6] = (double)S->arena_bytes;
    out[7] = (double)(S->ra_atoms.used() * kUnitBytes);
    out[8] = (double)S->ra_atoms.cap();
    out[9] = S->dev_s;
    return SMR_OK;
}

extern "C" int smr_server_destroy(smr_server *S) {
    if (!S) return SMR_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->ctx->device);
    S->ss.destroy();
    for (auto &kv : S->reqs) delete kv.second;
    free_host(S);
    smr_ctx *c = S->ctx;
    delete S;
    ctx_release(c);
    return SMR_OK;
}
