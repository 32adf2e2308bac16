// smr_internal.h -- host-side object layouts behind the opaque handles of
// surfmol_relax.h.  Shared by capi.cpp and server.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/surfmol_relax.h"
#include "smr_kernels.h"

#define SMR_MAX_ATOMS 1024

namespace smr {

extern thread_local std::string g_err;
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *where);

// owning device allocation
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release();
    hipError_t alloc(size_t nbytes);
    hipError_t upload(const void *src, size_t nbytes, hipStream_t st);
};

// scratch of the plain method's kernels for one system: hneigh [4N],
// force pieces [nf], per-atom bond/surface/pair energies [N], reduction
// output [16]
struct PlainScratch {
    float4 *hneigh, *fint;
    float *eatom, *esurf;
    double *epair, *cvf;
    int max_atoms, max_pieces;
};

struct HostPools {
    std::vector<TopoDesc> topo;
    std::vector<SysDesc> sys;
    std::vector<int4> ng;
    std::vector<float4> bk, bl, req, plq;
    std::vector<int2> a2f;
    std::vector<unsigned> ang_ng;
    std::vector<float4> ang_par;
    std::vector<uint2> dih_ng;
    std::vector<float4> dih_par;
    std::vector<uint2> inv_ng;
    std::vector<float4> inv_par;
    std::vector<unsigned short> a2f_idx;
    std::vector<uint2> ang_dst, dih_dst, inv_dst;
    // FP64 parameters for the energy evaluation
    std::vector<double4> bk64, bl64, req64, plq64, ang64, dih64, inv64;
};

struct DevPools {
    DevBuf topo, sys, ng, bk, bl, req, plq, a2f, ang_ng, ang_par, dih_ng, dih_par, inv_ng, inv_par, a2f_idx;
    DevBuf ang_dst, dih_dst, inv_dst;
    DevBuf bk64, bl64, req64, plq64, ang64, dih64, inv64;
};

// ---- size classes: systems are grouped by atom count; each class has its
//      own launch shape and its own persistent launch ----
constexpr int NCLASS = 5;
constexpr int kClassMaxAtoms[NCLASS] = {64, 128, 256, 512, 1024};
int size_class(int natoms);

struct ClassPlan {
    int cls = 0;
    int first = 0, count = 0;                 // slice of the work list
    int max_atoms = 0, nf_max = 0, na_max = 0, nd_max = 0, ni_max = 0;
    LaunchShape shape;
    size_t scratch_off = 0;                   // floats into the force-piece scratch
    int steps = 0;                            // FIRE steps of this launch (0: those of the launch arguments)
};

// launch-shape choice of one class remembered by the context (autotuner)
struct TunedShape {
    bool valid = false;
    int block = 0, apt = 0;
    bool fint_lds = false, pair_mode = true;
};

// shape with LDS bytes and persistent grid filled in; block == 0 if the
// combination does not exist or does not fit the device
LaunchShape make_shape(const smr_ctx *c, const ClassPlan &cp, int block, int apt, bool fint_lds, bool pair_mode,
                       bool stencil_cache, bool have_grid);
// default / tuned / overridden shape of a class
LaunchShape class_shape(const smr_ctx *c, const ClassPlan &cp, bool have_grid, int tune_block, int tune_cache,
                        int tune_pairs);
// group systems into classes; the work list is ordered by class (largest
// class first) and by falling atom count inside a class
struct SysSize {
    int id, natoms, nf, na, nd, ni;
};
void build_plans(std::vector<SysSize> sys, std::vector<ClassPlan> &plans, std::vector<int> &work);

// streams of one launcher (a context's batches, or a server's worker): the
// main stream plus three class streams, all at the default priority, and the
// fork/join events between them
struct StreamSet {
    hipStream_t main = nullptr;
    hipStream_t aux[3] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
    hipError_t create();
    void destroy();
};

// the launches of one relaxation chunk: counters zeroed by a kernel, one
// persistent launch per class, classes on separate streams (fork/join on the
// context stream), optionally captured once into a HIP graph
struct LaunchSet {
    DevBuf counters;              // int[NCLASS]
    DevBuf scratch;               // force pieces that do not fit LDS
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<unsigned char> key;
    bool use_graph = true;
    ~LaunchSet();
    void drop_graph();
};
int launch_plans(const StreamSet &ss, LaunchSet &ls, const std::vector<ClassPlan> &plans, const KernelArgs &base,
                 float *kernel_ms);
// the same without waiting: ev_begin / ev_end (may be null) are recorded on
// ss.main around the launches
int launch_plans_async(const StreamSet &ss, LaunchSet &ls, const std::vector<ClassPlan> &plans, const KernelArgs &base,
                       hipEvent_t ev_begin, hipEvent_t ev_end);
// the two halves of launch_plans, for callers that build their own sequence
int prepare_plans(LaunchSet &ls, const std::vector<ClassPlan> &plans, const KernelArgs &base,
                  std::vector<KernelArgs> &args);
hipError_t issue_plans(const StreamSet &ss, LaunchSet &ls, const std::vector<ClassPlan> &plans,
                       const std::vector<KernelArgs> &args,
                       bool zero_counters = true);

// ---- device-side start order (order_kernels.hip) ----
// launch slices of the work list: slice k covers positions [first[k], first[k+1])
struct RankSlices {
    int n = 0;
    int first[NCLASS + 1] = {0};
};
struct RankBuffers {
    DevBuf keys, vals, tmp;
    size_t tmp_bytes = 0;
    int n = 0;
};
hipError_t launch_reset_state(SysState *state, int n, double dt0, double alpha0, hipStream_t st);
// work[0..n) = work_base ranked by (slice, atom count desc, FP64 energy per
// atom of state[].E desc), stable; asynchronous on st
hipError_t rank_work_list(const DevTables &T, const SysState *state, const int *work_base, int n, const RankSlices &rs,
                          RankBuffers &rb, int *work, hipStream_t st);

// page-locked host memory (released with the owner)
struct HostPinned {
    void *ptr = nullptr;
    size_t bytes = 0;
    HostPinned() = default;
    HostPinned(const HostPinned &) = delete;
    HostPinned &operator=(const HostPinned &) = delete;
    ~HostPinned() { if (ptr) (void)hipHostFree(ptr); }
    hipError_t alloc(size_t nbytes) {
        if (ptr) { (void)hipHostFree(ptr); ptr = nullptr; bytes = 0; }
        hipError_t e = hipHostMalloc(&ptr, nbytes, hipHostMallocDefault);
        if (e == hipSuccess) bytes = nbytes; else ptr = nullptr;
        return e;
    }
};

int pack_topology(const smr_topology &t, int ti, double alpha, bool have_grid, HostPools &P, TopoDesc &td);
hipError_t upload_pools(HostPools &P, DevPools &D, hipStream_t st);
void fill_tables(const DevPools &D, const smr_gridff *grid, const smr_ff_options &o, DevTables &T);
void system_origin(const double *x, int n, const smr_gridff *grid, bool surface, double org[3], int &iux, int &iuy,
                   int g[3]);
int check_fire(const smr_fire_params *f, const char *who);
void ctx_retain(smr_ctx *c);
void ctx_release(smr_ctx *c);
FireConst make_fire_const(const smr_fire_params *f, int max_steps, double fconv);

}  // namespace smr

struct smr_ctx {
    int refs = 1;                // the caller's handle + one per live grid/batch/server
    bool destroy_requested = false;
    int device = 0;
    int num_cu = 0;
    size_t max_lds = 0;
    std::string arch;
    hipStream_t stream = nullptr;   // == ss.main
    smr::StreamSet ss;              // used by the context's batches (one caller thread)
    smr::TunedShape tuned[smr::NCLASS];
};

struct smr_gridff {
    smr_ctx *ctx = nullptr;
    smr_grid_desc desc;
    smr::DevBuf coef;        // float4 [nx*ny*nz]
    smr::DevBuf reference;   // double4 [nx*ny*nz] (optional)
};

struct smr_batch {
    smr_ctx *ctx = nullptr;
    const smr_gridff *grid = nullptr;
    smr_ff_options opts;
    smr::HostPools hp;
    smr::DevPools dp;
    DevTables tab;
    int nsys = 0, total_atoms = 0, max_atoms = 0, nf_max = 0, na_max = 0, nd_max = 0, ni_max = 0;
    std::vector<int> atom_off;       // [nsys+1]
    std::vector<double> origin;      // [3*nsys]
    smr::DevBuf pos, pos0, vel, force, state, work;
    smr::DevBuf stencil;             // float4 [16 * total_atoms] stored surface stencils
    smr::DevBuf stencil_key;         // int [total_atoms] cell of each stored stencil
    // coupled layer (opts.inter_cutoff > 0)
    bool layer = false;
    smr::DevBuf atom_sys, req_atom, cell_count, cell_start, cell_sidx, cell_spos, cell_sreq, atom_cell;
    smr::DevBuf atom_wrap, cell_swrap;
    smr::DevBuf cell_part, slot_of;
    smr::DevBuf partials, layer_state;
    smr::CellGrid cells;
    // coupled layer in one cooperative launch (allocated on first use)
    smr::DevBuf lp_xbuf, lp_xrec, lp_bar;
    int layer_persistent = 1;
    float layer_skin = 1.f;
    int last_layer_persistent = 0, last_layer_fallback = 0, last_layer_rebuilds = 0;
    smr::DevBuf plain_hneigh, plain_fint, plain_e, plain_cvf;
    std::vector<smr::ClassPlan> class_plans;   // one per size class (shapes chosen per class)
    std::vector<smr::ClassPlan> plans;         // launches: adjacent classes of one shape merged
    smr::LaunchSet ls;
    smr::LaunchSet ls_rerank;        // the continuation launch of a re-ranked relaxation
    smr::DevBuf work_base;           // work_host on the device: the ranking's input
    smr::RankBuffers rank;           // keys, values and sort scratch of the ranking
    smr::HostPinned hstate;          // SysState [nsys]: results read back here
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // launch timing (created on first use)
    ~smr_batch();
    int tune_block = 0;              // 0: automatic
    int tune_cache = -1;             // -1: automatic
    int tune_pairs = -1;             // -1: automatic
    int energy_order = 1;            // 1: start systems of higher initial energy per atom first
    int rerank_steps = -1;           // > 0: re-rank the unfinished systems after this many steps;
                                     // -1: after max_steps / 5; 0: never
    std::vector<int> work_host;      // the work list as uploaded (class order, atom count inside a class)
    float last_kernel_ms = 0.f;
};

namespace smr {
// the plain method (capi.cpp): scratch from `ext`, or the batch's own
// buffers (allocated on first use) when ext == nullptr
int relax_plain_run(smr_batch *b, int isys, const smr_fire_params *fire, int max_steps, double fconv,
                    smr_relax_result *result, const PlainScratch *ext, const char *who);
// the fixed scratch of smr_relax_plain_noalloc on `device` (current):
// resolves its device address once per device (plain_noalloc.hip)
int noalloc_setup(int device);
}  // namespace smr
