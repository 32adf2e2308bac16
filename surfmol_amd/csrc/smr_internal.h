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
};

struct DevPools {
    DevBuf topo, sys, ng, bk, bl, req, plq, a2f, ang_ng, ang_par, dih_ng, dih_par, inv_ng, inv_par, a2f_idx;
};

int pack_topology(const smr_topology &t, int ti, double alpha, bool have_grid, HostPools &P, TopoDesc &td);
hipError_t upload_pools(HostPools &P, DevPools &D, hipStream_t st);
void system_origin(const double *x, int n, const smr_gridff *grid, bool surface, double org[3], int &iux, int &iuy);
LaunchShape choose_shape(const smr_ctx *c, const smr_batch *b, int nwork);
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
    hipStream_t stream = nullptr;
    smr::DevBuf counter;
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
    smr::DevBuf pos, pos0, vel, force, state, work, fint_scratch;
    smr::DevBuf plain_hneigh, plain_fint, plain_e, plain_cvf;
    smr::LaunchShape last_shape;
    int tune_block = 0;              // 0: automatic
    int tune_cache = -1;             // -1: automatic
    float last_kernel_ms = 0.f;
};
