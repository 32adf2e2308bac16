// capi.cpp -- implementation of the C ABI in include/surfmol_relax.h.
//
// Host side of the device boundary: validates and packs the caller's
// Uff/NonBondedFF tables into the device pools of smr_device.h, owns device
// memory and streams, and drives the kernels of relax_kernels.hip and
// gridff_kernels.hip.  No device code here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/surfmol_relax.h"
#include "smr_internal.h"
#include "smr_kernels.h"

using namespace smr;

namespace smr {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *where) {
    return fail(SMR_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

hipError_t DevBuf::alloc(size_t nbytes) {
    release();
    if (nbytes == 0) nbytes = 16;
    hipError_t e = hipMalloc(&ptr, nbytes);
    if (e == hipSuccess) bytes = nbytes;
    else ptr = nullptr;
    return e;
}

hipError_t DevBuf::upload(const void *src, size_t nbytes, hipStream_t st) {
    hipError_t e = alloc(nbytes);
    if (e != hipSuccess) return e;
    if (nbytes == 0) return hipSuccess;
    return hipMemcpyAsync(ptr, src, nbytes, hipMemcpyHostToDevice, st);
}
}  // namespace smr

#define HIPCHK(expr)                                              \
    do {                                                          \
        hipError_t _e = (expr);                                   \
        if (_e != hipSuccess) return hip_fail(_e, #expr);         \
    } while (0)

extern "C" const char *smr_last_error(void) { return g_err.c_str(); }
extern "C" const char *smr_version(void) { return "surfmol_relax 0.1 (gfx950)"; }

// ============================================================ context
extern "C" int smr_ctx_create(int device, smr_ctx **out) {
    if (!out) return fail(SMR_ERR_INVALID, "smr_ctx_create: out == NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return fail(SMR_ERR_HIP, std::string("smr_ctx_create: no HIP device (") + hipGetErrorString(e) + ")");
    if (device < 0 || device >= ndev)
        return fail(SMR_ERR_INVALID, "smr_ctx_create: device " + std::to_string(device) + " not in [0," +
                                         std::to_string(ndev) + ")");
    HIPCHK(hipSetDevice(device));
    auto *c = new smr_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) { delete c; return hip_fail(e, "hipGetDeviceProperties"); }
    c->num_cu = prop.multiProcessorCount;
    c->max_lds = prop.sharedMemPerBlock;
    c->arch = prop.gcnArchName;
    e = c->ss.create();
    if (e != hipSuccess) { delete c; return hip_fail(e, "smr_ctx_create: streams"); }
    int rc = noalloc_setup(device);
    if (rc) { c->ss.destroy(); delete c; return rc; }
    c->stream = c->ss.main;
    *out = c;
    return SMR_OK;
}

namespace smr {
// A context lives until its handle has been destroyed and every grid, batch
// and server made from it is gone, whatever the order of the destroy calls.
void ctx_retain(smr_ctx *c) { c->refs++; }

void ctx_release(smr_ctx *c) {
    if (--c->refs > 0) return;
    (void)hipSetDevice(c->device);
    c->ss.destroy();
    delete c;
}
}  // namespace smr

extern "C" int smr_ctx_destroy(smr_ctx *c) {
    if (!c) return SMR_OK;
    if (c->destroy_requested) return fail(SMR_ERR_STATE, "smr_ctx_destroy: context already destroyed");
    c->destroy_requested = true;
    ctx_release(c);
    return SMR_OK;
}

extern "C" int smr_ctx_info(smr_ctx *c, double *out) {
    if (!c || !out) return fail(SMR_ERR_INVALID, "smr_ctx_info: NULL argument");
    HIPCHK(hipSetDevice(c->device));
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    int ws = 0;
    HIPCHK(hipDeviceGetAttribute(&ws, hipDeviceAttributeWarpSize, c->device));
    out[0] = (double)fr;
    out[1] = (double)tot;
    out[2] = (double)c->num_cu;
    out[3] = (double)c->max_lds;
    out[4] = (double)ws;
    return SMR_OK;
}

// ============================================================ GridFF
static int check_desc(const smr_grid_desc *d, const char *who) {
    if (!d) return fail(SMR_ERR_INVALID, std::string(who) + ": desc == NULL");
    if (d->nx < 4 || d->ny < 4 || d->nz < 4)
        return fail(SMR_ERR_INVALID, std::string(who) + ": grid " + std::to_string(d->nx) + "x" + std::to_string(d->ny) +
                                         "x" + std::to_string(d->nz) + " too small (>= 4 per axis)");
    for (int k = 0; k < 3; k++)
        if (!(d->step[k] > 0.0))
            return fail(SMR_ERR_INVALID, std::string(who) + ": step[" + std::to_string(k) + "] must be > 0");
    if ((double)d->nx * d->ny * d->nz > 4.0e9)
        return fail(SMR_ERR_UNSUPPORTED, std::string(who) + ": more than 4e9 nodes");
    return SMR_OK;
}

extern "C" int smr_gridff_upload(smr_ctx *c, const smr_grid_desc *d, const float *coefs, smr_gridff **out) {
    if (!c || !out || !coefs) return fail(SMR_ERR_INVALID, "smr_gridff_upload: NULL argument");
    *out = nullptr;
    int rc = check_desc(d, "smr_gridff_upload");
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    auto *g = new smr_gridff();
    g->ctx = c;
    ctx_retain(c);
    g->desc = *d;
    size_t n = (size_t)d->nx * d->ny * d->nz;
    hipError_t e = g->coef.upload(coefs, n * 4 * sizeof(float), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { delete g; ctx_release(c); return hip_fail(e, "smr_gridff_upload"); }
    *out = g;
    return SMR_OK;
}

extern "C" int smr_gridff_build(smr_ctx *c, const smr_grid_desc *d, int nsub, const double *sub_pos,
                                const double *sub_reqs, int npbc, smr_gridff **out) {
    if (!c || !out) return fail(SMR_ERR_INVALID, "smr_gridff_build: NULL argument");
    *out = nullptr;
    int rc = check_desc(d, "smr_gridff_build");
    if (rc) return rc;
    if (nsub <= 0 || !sub_pos || !sub_reqs) return fail(SMR_ERR_INVALID, "smr_gridff_build: need nsub > 0 substrate atoms");
    if (npbc < 0 || npbc > 64) return fail(SMR_ERR_INVALID, "smr_gridff_build: npbc " + std::to_string(npbc) + " not in [0,64]");
    if (d->nx < 3 || d->ny < 3) return fail(SMR_ERR_INVALID, "smr_gridff_build: periodic axes need >= 3 nodes");
    HIPCHK(hipSetDevice(c->device));
    auto *g = new smr_gridff();
    g->ctx = c;
    ctx_retain(c);
    g->desc = *d;
    const size_t n = (size_t)d->nx * d->ny * d->nz;
    DevBuf dpos, dreq, E;
    hipError_t e = dpos.upload(sub_pos, sizeof(double) * 3 * nsub, c->stream);
    if (e == hipSuccess) e = dreq.upload(sub_reqs, sizeof(double) * 4 * nsub, c->stream);
    if (e == hipSuccess) e = E.alloc(sizeof(double) * 4 * n);
    if (e != hipSuccess) { delete g; ctx_release(c); return hip_fail(e, "smr_gridff_build: allocation"); }
    GridBuildArgs G;
    G.nx = d->nx; G.ny = d->ny; G.nz = d->nz;
    for (int k = 0; k < 3; k++) G.p0[k] = d->p0[k];
    G.L[0] = d->nx * d->step[0];
    G.L[1] = d->ny * d->step[1];
    G.L[2] = d->nz * d->step[2];
    G.npbc = npbc;
    G.alpha = d->alpha_morse;
    G.r2damp = d->rdamp * d->rdamp;
    G.nsub = nsub;
    G.sub_pos = (const double *)dpos.ptr;
    G.sub_req = (const double *)dreq.ptr;
    G.out = (double *)E.ptr;
    e = launch_gridff_build(G, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { delete g; ctx_release(c); return hip_fail(e, "smr_gridff_build: build kernel"); }
    if (d->keep_reference) {
        e = g->reference.alloc(sizeof(double) * 4 * n);
        if (e == hipSuccess)
            e = hipMemcpyAsync(g->reference.ptr, E.ptr, sizeof(double) * 4 * n, hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) { delete g; ctx_release(c); return hip_fail(e, "smr_gridff_build: keep_reference"); }
    }
    e = launch_bspline_fit((double *)E.ptr, d->nx, d->ny, d->nz, nullptr, c->stream);
    if (e != hipSuccess) { delete g; ctx_release(c); return hip_fail(e, "smr_gridff_build: B-spline fit"); }
    e = g->coef.alloc(sizeof(float) * 4 * n);
    if (e == hipSuccess) e = launch_pack_f32((const double *)E.ptr, (float *)g->coef.ptr, n, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { delete g; ctx_release(c); return hip_fail(e, "smr_gridff_build: pack"); }
    *out = g;
    return SMR_OK;
}

extern "C" int smr_gridff_download(const smr_gridff *g, float *coefs) {
    if (!g || !coefs) return fail(SMR_ERR_INVALID, "smr_gridff_download: NULL argument");
    HIPCHK(hipSetDevice(g->ctx->device));
    HIPCHK(hipMemcpy(coefs, g->coef.ptr, g->coef.bytes, hipMemcpyDeviceToHost));
    return SMR_OK;
}

// ---- on-disk coefficient cache: the reference's lazy file cache of fitted
//      grids (GridFF.h tryLoad_new loads Bspline_PLQd.npy and refits only
//      when it is missing, notes/designs/2026-08-29_gridff_faf_porting_notes.md:
//      393-394).  Files are .npy, float32, shape (nx, ny, nz, 4): the device
//      node layout iz + nz*(iy + ny*ix) in C order. ----
namespace {

std::string npy_header(const smr_grid_desc *d) {
    std::string h = "{'descr': '<f4', 'fortran_order': False, 'shape': (" + std::to_string(d->nx) + ", " +
                    std::to_string(d->ny) + ", " + std::to_string(d->nz) + ", 4), }";
    // magic (6) + version (2) + length (2) + header + '\n' is a multiple of 64
    const size_t total = 10 + h.size() + 1;
    h.append((64 - total % 64) % 64, ' ');
    h += '\n';
    return h;
}

// FNV-1a over the bytes of every build input
struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void *p, size_t n) {
        const unsigned char *b = (const unsigned char *)p;
        for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ull; }
    }
    template <class T> void val(const T &v) { add(&v, sizeof(T)); }
};

}  // namespace

extern "C" int smr_gridff_save(const smr_gridff *g, const char *path) {
    if (!g || !path) return fail(SMR_ERR_INVALID, "smr_gridff_save: NULL argument");
    HIPCHK(hipSetDevice(g->ctx->device));
    std::vector<float> h(g->coef.bytes / sizeof(float));
    HIPCHK(hipMemcpy(h.data(), g->coef.ptr, g->coef.bytes, hipMemcpyDeviceToHost));
    const std::string hdr = npy_header(&g->desc);
    // written to a temporary name, then renamed: a reader never sees half a file
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return fail(SMR_ERR_INVALID, std::string("smr_gridff_save: cannot open ") + tmp);
        const unsigned char pre[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
        f.write((const char *)pre, 8);
        const uint16_t hl = (uint16_t)hdr.size();
        const unsigned char len[2] = {(unsigned char)(hl & 0xff), (unsigned char)(hl >> 8)};
        f.write((const char *)len, 2);
        f.write(hdr.data(), (std::streamsize)hdr.size());
        f.write((const char *)h.data(), (std::streamsize)(h.size() * sizeof(float)));
        if (!f) return fail(SMR_ERR_INVALID, std::string("smr_gridff_save: write failed: ") + tmp);
    }
    if (std::rename(tmp.c_str(), path) != 0) return fail(SMR_ERR_INVALID, std::string("smr_gridff_save: rename to ") + path);
    return SMR_OK;
}

extern "C" int smr_gridff_load(smr_ctx *c, const smr_grid_desc *d, const char *path, smr_gridff **out) {
    if (!c || !d || !path || !out) return fail(SMR_ERR_INVALID, "smr_gridff_load: NULL argument");
    *out = nullptr;
    int rc = check_desc(d, "smr_gridff_load");
    if (rc) return rc;
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(SMR_ERR_INVALID, std::string("smr_gridff_load: cannot open ") + path);
    unsigned char pre[10];
    f.read((char *)pre, 10);
    if (!f || std::memcmp(pre, "\x93NUMPY", 6) != 0 || pre[6] != 1)
        return fail(SMR_ERR_INVALID, std::string("smr_gridff_load: ") + path + " is not a version-1 .npy file");
    const size_t hl = (size_t)pre[8] | ((size_t)pre[9] << 8);
    std::string hdr(hl, ' ');
    f.read(&hdr[0], (std::streamsize)hl);
    if (!f || hdr != npy_header(d))
        return fail(SMR_ERR_INVALID, std::string("smr_gridff_load: ") + path + " holds another dtype or shape than (" +
                                         std::to_string(d->nx) + ", " + std::to_string(d->ny) + ", " +
                                         std::to_string(d->nz) + ", 4) float32");
    const size_t n = (size_t)d->nx * d->ny * d->nz * 4;
    std::vector<float> h(n);
    f.read((char *)h.data(), (std::streamsize)(n * sizeof(float)));
    if (!f || f.peek() != std::char_traits<char>::eof())
        return fail(SMR_ERR_INVALID, std::string("smr_gridff_load: ") + path + " has the wrong payload size");
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(h[i])) return fail(SMR_ERR_INVALID, std::string("smr_gridff_load: non-finite value in ") + path);
    return smr_gridff_upload(c, d, h.data(), out);
}

extern "C" int smr_gridff_build_cached(smr_ctx *c, const smr_grid_desc *d, int nsub, const double *sub_pos,
                                       const double *sub_reqs, int npbc, const char *cache_dir, int *hit,
                                       smr_gridff **out) {
    if (!c || !d || !out || !cache_dir) return fail(SMR_ERR_INVALID, "smr_gridff_build_cached: NULL argument");
    *out = nullptr;
    if (hit) *hit = 0;
    if (nsub <= 0 || !sub_pos || !sub_reqs) return fail(SMR_ERR_INVALID, "smr_gridff_build_cached: need nsub > 0 substrate atoms");
    Fnv k;
    const char tag[] = "smr-gridff-v1";
    k.add(tag, sizeof(tag));
    k.val(d->nx); k.val(d->ny); k.val(d->nz);
    k.add(d->p0, sizeof(d->p0)); k.add(d->step, sizeof(d->step));
    k.val(d->alpha_morse); k.val(d->rdamp); k.val(npbc); k.val(nsub);
    k.add(sub_pos, sizeof(double) * 3 * (size_t)nsub);
    k.add(sub_reqs, sizeof(double) * 4 * (size_t)nsub);
    char name[40];
    std::snprintf(name, sizeof(name), "gridff_%016llx.npy", (unsigned long long)k.h);
    const std::string path = std::string(cache_dir) + "/" + name;
    // the FP64 reference values are not cached: a grid that keeps them is built
    if (!d->keep_reference) {
        std::ifstream probe(path, std::ios::binary);
        if (probe.good()) {
            probe.close();
            if (smr_gridff_load(c, d, path.c_str(), out) == SMR_OK) {
                if (hit) *hit = 1;
                return SMR_OK;
            }
            // an unusable entry is rebuilt and overwritten
        }
    }
    int rc = smr_gridff_build(c, d, nsub, sub_pos, sub_reqs, npbc, out);
    if (rc) return rc;
    rc = smr_gridff_save(*out, path.c_str());
    if (rc) {
        (void)smr_gridff_destroy(*out);
        *out = nullptr;
        return rc;
    }
    return SMR_OK;
}

extern "C" int smr_gridff_download_reference(const smr_gridff *g, double *values) {
    if (!g || !values) return fail(SMR_ERR_INVALID, "smr_gridff_download_reference: NULL argument");
    if (!g->reference.ptr)
        return fail(SMR_ERR_STATE, "smr_gridff_download_reference: grid was not built with keep_reference = 1");
    HIPCHK(hipSetDevice(g->ctx->device));
    HIPCHK(hipMemcpy(values, g->reference.ptr, g->reference.bytes, hipMemcpyDeviceToHost));
    return SMR_OK;
}

extern "C" int smr_gridff_destroy(smr_gridff *g) {
    if (!g) return SMR_OK;
    smr_ctx *c = g->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete g;
    ctx_release(c);
    return SMR_OK;
}

// ============================================================ batch packing
namespace smr {

int pack_topology(const smr_topology &t, int ti, double alpha, bool have_grid, HostPools &P, TopoDesc &td) {
    const std::string who = "topology " + std::to_string(ti) + ": ";
    const int N = t.natoms;
    if (N <= 0) return fail(SMR_ERR_INVALID, who + "natoms must be > 0");
    if (N > SMR_MAX_ATOMS)
        return fail(SMR_ERR_UNSUPPORTED, who + "natoms " + std::to_string(N) + " > " + std::to_string(SMR_MAX_ATOMS));
    if (t.nbonds < 0 || t.nangles < 0 || t.ndih < 0 || t.ninv < 0) return fail(SMR_ERR_INVALID, who + "negative counts");
    if (!t.neighs || !t.neigh_bs || !t.reqs || (t.nbonds && !t.bon_params) ||
        (t.nangles && (!t.ang_atoms || !t.ang_ngs || !t.ang_params)) ||
        (t.ndih && (!t.dih_atoms || !t.dih_ngs || !t.dih_params)) || (t.ninv && (!t.inv_atoms || !t.inv_ngs || !t.inv_params)))
        return fail(SMR_ERR_INVALID, who + "NULL table for a non-empty term list");
    td.natoms = N;
    td.nangles = t.nangles;
    td.ndih = t.ndih;
    td.ninv = t.ninv;
    td.nf = 4 * t.ndih + 4 * t.ninv + 3 * t.nangles;
    if (td.nf > 65535 || 4 * N > 65535)
        return fail(SMR_ERR_UNSUPPORTED, who + std::to_string(td.nf) + " force pieces; 16-bit gather lists hold at most 65535");
    td.atom_off = (int)P.ng.size();
    td.ang_off = (int)P.ang_ng.size();
    td.dih_off = (int)P.dih_ng.size();
    td.inv_off = (int)P.inv_ng.size();
    td.a2f_off = (int)P.a2f_idx.size();
    // ---- atoms ----
    for (int i = 0; i < N; i++) {
        int ng[4];
        float bk[4] = {0, 0, 0, 0}, bl[4] = {0, 0, 0, 0};
        for (int s = 0; s < 4; s++) {
            ng[s] = t.neighs[4 * i + s];
            if (ng[s] < -1 || ng[s] >= N || ng[s] == i)
                return fail(SMR_ERR_INVALID, who + "neighs[" + std::to_string(i) + "][" + std::to_string(s) + "] = " +
                                                 std::to_string(ng[s]) + " out of range");
            if (ng[s] >= 0) {
                int ib = t.neigh_bs[4 * i + s];
                if (ib < 0 || ib >= t.nbonds)
                    return fail(SMR_ERR_INVALID, who + "neigh_bs[" + std::to_string(i) + "][" + std::to_string(s) + "] = " +
                                                     std::to_string(ib) + " not a bond index");
                bk[s] = (float)t.bon_params[2 * ib];
                bl[s] = (float)t.bon_params[2 * ib + 1];
            }
        }
        P.ng.push_back(make_int4(ng[0], ng[1], ng[2], ng[3]));
        P.bk.push_back(make_float4(bk[0], bk[1], bk[2], bk[3]));
        P.bl.push_back(make_float4(bl[0], bl[1], bl[2], bl[3]));
        double bk64[4] = {0, 0, 0, 0}, bl64[4] = {0, 0, 0, 0};
        for (int s = 0; s < 4; s++)
            if (ng[s] >= 0) {
                bk64[s] = t.bon_params[2 * t.neigh_bs[4 * i + s]];
                bl64[s] = t.bon_params[2 * t.neigh_bs[4 * i + s] + 1];
            }
        P.bk64.push_back(make_double4(bk64[0], bk64[1], bk64[2], bk64[3]));
        P.bl64.push_back(make_double4(bl64[0], bl64[1], bl64[2], bl64[3]));
        const double *r = t.reqs + 4 * i;
        // pair-term row: (R, sqrtE, sqrt(kC)*Q) so that qi*qj = kC*Qi*Qj
        const double sq = r[2] * std::sqrt(14.3996448915);
        P.req.push_back(make_float4((float)r[0], (float)r[1], (float)sq, 0.f));
        P.req64.push_back(make_double4(r[0], r[1], sq, 0.0));
        // PLQH (opencl/UFF.cl:1681-1687), computed in FP64
        if (have_grid) {
            double ej = std::exp(alpha * r[0]);
            P.plq.push_back(make_float4((float)(ej * ej * r[1]), (float)(ej * r[1]), (float)r[2], 0.f));
            P.plq64.push_back(make_double4(ej * ej * r[1], ej * r[1], r[2], 0.0));
        } else {
            P.plq.push_back(make_float4(0.f, 0.f, 0.f, 0.f));
            P.plq64.push_back(make_double4(0.0, 0.0, 0.0, 0.0));
        }
    }
    auto slot_ok = [&](int g) { return g >= 0 && g < 4 * N && t.neighs[g] >= 0; };
    // hneigh slot 4*atom + k of the reference -> k*N + atom on the device
    // (slot-major, so that the per-atom bond phase writes consecutive rows)
    auto dslot = [&](int g) { return (unsigned)((g & 3) * N + (g >> 2)); };
    // ---- angles ----
    for (int a = 0; a < t.nangles; a++) {
        int g0 = t.ang_ngs[2 * a], g1 = t.ang_ngs[2 * a + 1];
        if (!slot_ok(g0) || !slot_ok(g1))
            return fail(SMR_ERR_INVALID, who + "angle " + std::to_string(a) + " has hneigh slots (" + std::to_string(g0) +
                                             "," + std::to_string(g1) + ") that are not bonds");
        const double *p = t.ang_params + 5 * a;
        P.ang_ng.push_back(dslot(g0) | (dslot(g1) << 16));
        P.ang_par.push_back(make_float4((float)(p[0] * p[1]), (float)(p[0] * p[2]), (float)(p[0] * p[3]), (float)(p[0] * p[4])));
        P.ang64.push_back(make_double4(p[0], p[1], p[3], p[4]));
    }
    // ---- dihedrals ----
    for (int d = 0; d < t.ndih; d++) {
        int g0 = t.dih_ngs[3 * d], g1 = t.dih_ngs[3 * d + 1], g2 = t.dih_ngs[3 * d + 2];
        if (!slot_ok(g0) || !slot_ok(g1) || !slot_ok(g2))
            return fail(SMR_ERR_INVALID, who + "dihedral " + std::to_string(d) + " has an invalid hneigh slot");
        const double *p = t.dih_params + 3 * d;
        int n = (int)p[2];
        if (n < 1 || n > 12) return fail(SMR_ERR_INVALID, who + "dihedral " + std::to_string(d) + " multiplicity " + std::to_string(n));
        P.dih_ng.push_back(make_uint2(dslot(g0) | (dslot(g1) << 16), dslot(g2) | ((unsigned)n << 16)));
        P.dih_par.push_back(make_float4((float)p[0], (float)(p[0] * p[1]), (float)(p[0] * p[1] * p[2]), 0.f));
        P.dih64.push_back(make_double4(p[0], p[1], (double)n, 0.0));
    }
    // ---- inversions ----
    for (int v = 0; v < t.ninv; v++) {
        int g0 = t.inv_ngs[3 * v], g1 = t.inv_ngs[3 * v + 1], g2 = t.inv_ngs[3 * v + 2];
        if (!slot_ok(g0) || !slot_ok(g1) || !slot_ok(g2))
            return fail(SMR_ERR_INVALID, who + "inversion " + std::to_string(v) + " has an invalid hneigh slot");
        const double *p = t.inv_params + 4 * v;
        P.inv_ng.push_back(make_uint2(dslot(g0) | (dslot(g1) << 16), dslot(g2)));
        P.inv_par.push_back(make_float4((float)(p[0] * p[1]), (float)(p[0] * p[2]), (float)(p[0] * p[3]), 0.f));
        P.inv64.push_back(make_double4(p[0], p[1], p[2], p[3]));
    }
    // ---- a2f gather lists (Uff::map_atom_interactions, uff.rs:317-370):
    //      dihedral pieces, then inversion pieces, then angle pieces.  The
    //      device stores piece k of term t at k*nterms + t within each term
    //      block (consecutive terms write consecutive slots); the gather
    //      order per atom is the reference's. ----
    std::vector<int> cnt(N, 0), off(N + 1, 0);
    auto chk_atom = [&](int a) { return a >= 0 && a < N; };
    for (int d = 0; d < t.ndih; d++)
        for (int k = 0; k < 4; k++) {
            int a = t.dih_atoms[4 * d + k];
            if (!chk_atom(a)) return fail(SMR_ERR_INVALID, who + "dihedral atom out of range");
            cnt[a]++;
        }
    for (int v = 0; v < t.ninv; v++)
        for (int k = 0; k < 4; k++) {
            int a = t.inv_atoms[4 * v + k];
            if (!chk_atom(a)) return fail(SMR_ERR_INVALID, who + "inversion atom out of range");
            cnt[a]++;
        }
    for (int a3 = 0; a3 < t.nangles; a3++)
        for (int k = 0; k < 3; k++) {
            int a = t.ang_atoms[3 * a3 + k];
            if (!chk_atom(a)) return fail(SMR_ERR_INVALID, who + "angle atom out of range");
            cnt[a]++;
        }
    for (int i = 0; i < N; i++) off[i + 1] = off[i] + cnt[i];
    std::vector<unsigned short> idx(off[N]);
    std::fill(cnt.begin(), cnt.end(), 0);
    const int i0inv = 4 * t.ndih, i0ang = i0inv + 4 * t.ninv;
    for (int d = 0; d < t.ndih; d++)
        for (int k = 0; k < 4; k++) { int a = t.dih_atoms[4 * d + k]; idx[off[a] + cnt[a]++] = (unsigned short)(k * t.ndih + d); }
    for (int v = 0; v < t.ninv; v++)
        for (int k = 0; k < 4; k++) { int a = t.inv_atoms[4 * v + k]; idx[off[a] + cnt[a]++] = (unsigned short)(i0inv + k * t.ninv + v); }
    for (int a3 = 0; a3 < t.nangles; a3++)
        for (int k = 0; k < 3; k++) { int a = t.ang_atoms[3 * a3 + k]; idx[off[a] + cnt[a]++] = (unsigned short)(i0ang + k * t.nangles + a3); }
    for (int i = 0; i < N; i++) P.a2f.push_back(make_int2(off[i], off[i + 1] - off[i]));
    P.a2f_idx.insert(P.a2f_idx.end(), idx.begin(), idx.end());
    // the inverse: the atom-major position of every piece, per term
    std::vector<unsigned> dst(td.nf, 0);
    for (int q = 0; q < off[N]; q++) dst[idx[q]] = (unsigned)q;
    auto pk = [&](int p0, int p1) { return dst[p0] | (dst[p1] << 16); };
    for (int d = 0; d < t.ndih; d++) {
        const int n = t.ndih;
        P.dih_dst.push_back(make_uint2(pk(d, n + d), pk(2 * n + d, 3 * n + d)));
    }
    for (int v = 0; v < t.ninv; v++) {
        const int n = t.ninv;
        P.inv_dst.push_back(make_uint2(pk(i0inv + v, i0inv + n + v), pk(i0inv + 2 * n + v, i0inv + 3 * n + v)));
    }
    for (int a3 = 0; a3 < t.nangles; a3++) {
        const int n = t.nangles;
        P.ang_dst.push_back(make_uint2(pk(i0ang + a3, i0ang + n + a3), dst[i0ang + 2 * n + a3]));
    }
    return SMR_OK;
}

hipError_t upload_pools(HostPools &P, DevPools &D, hipStream_t st) {
    hipError_t e;
#define UP(name) if ((e = D.name.upload(P.name.data(), P.name.size() * sizeof(P.name[0]), st)) != hipSuccess) return e;
    UP(topo) UP(sys) UP(ng) UP(bk) UP(bl) UP(req) UP(plq) UP(a2f) UP(ang_ng) UP(ang_par) UP(dih_ng) UP(dih_par)
    UP(inv_ng) UP(inv_par) UP(a2f_idx) UP(ang_dst) UP(dih_dst) UP(inv_dst) UP(bk64) UP(bl64) UP(req64) UP(plq64) UP(ang64) UP(dih64) UP(inv64)
#undef UP
    return hipSuccess;
}

void fill_tables(const DevPools &D, const smr_gridff *grid, const smr_ff_options &o, DevTables &T) {
    T.topo = (const TopoDesc *)D.topo.ptr;
    T.sys = (const SysDesc *)D.sys.ptr;
    T.ng = (const int4 *)D.ng.ptr;
    T.bk = (const float4 *)D.bk.ptr;
    T.bl = (const float4 *)D.bl.ptr;
    T.req = (const float4 *)D.req.ptr;
    T.plq = (const float4 *)D.plq.ptr;
    T.a2f = (const int2 *)D.a2f.ptr;
    T.ang_ng = (const unsigned *)D.ang_ng.ptr;
    T.ang_par = (const float4 *)D.ang_par.ptr;
    T.dih_ng = (const uint2 *)D.dih_ng.ptr;
    T.dih_par = (const float4 *)D.dih_par.ptr;
    T.inv_ng = (const uint2 *)D.inv_ng.ptr;
    T.inv_par = (const float4 *)D.inv_par.ptr;
    T.a2f_idx = (const unsigned short *)D.a2f_idx.ptr;
    T.ang_dst = (const uint2 *)D.ang_dst.ptr;
    T.dih_dst = (const uint2 *)D.dih_dst.ptr;
    T.inv_dst = (const uint2 *)D.inv_dst.ptr;
    T.bk64 = (const double4 *)D.bk64.ptr;
    T.bl64 = (const double4 *)D.bl64.ptr;
    T.req64 = (const double4 *)D.req64.ptr;
    T.plq64 = (const double4 *)D.plq64.ptr;
    T.ang64 = (const double4 *)D.ang64.ptr;
    T.dih64 = (const double4 *)D.dih64.ptr;
    T.inv64 = (const double4 *)D.inv64.ptr;
    T.r2damp64 = o.rdamp * o.rdamp;
    if (grid && o.surface) {
        T.grid = (const float4 *)grid->coef.ptr;
        T.gnx = grid->desc.nx; T.gny = grid->desc.ny; T.gnz = grid->desc.nz;
        T.ginv_x = (float)(1.0 / grid->desc.step[0]);
        T.ginv_y = (float)(1.0 / grid->desc.step[1]);
        T.ginv_z = (float)(1.0 / grid->desc.step[2]);
        T.ginv64_x = 1.0 / grid->desc.step[0];
        T.ginv64_y = 1.0 / grid->desc.step[1];
        T.ginv64_z = 1.0 / grid->desc.step[2];
        T.ostep64_x = grid->desc.step[0];
        T.ostep64_y = grid->desc.step[1];
        T.ostep64_z = grid->desc.step[2];
    } else {
        T.grid = nullptr;
        T.gnx = T.gny = T.gnz = 1;
        T.ginv_x = T.ginv_y = T.ginv_z = 1.f;
        T.ginv64_x = T.ginv64_y = T.ginv64_z = 1.0;
        T.ostep64_x = T.ostep64_y = T.ostep64_z = 1.0;
    }
    T.ostep_x = (float)T.ostep64_x;
    T.ostep_y = (float)T.ostep64_y;
    T.ostep_z = (float)T.ostep64_z;
    T.r2damp = (float)(o.rdamp * o.rdamp);
    T.use_pairs = o.nonbonded ? 1 : 0;
}

// system origin: grid-node aligned in x,y (so the surface lookup offset is an
// exact integer), the grid's z origin in z; without a grid, the centroid
// rounded to whole Angstroms.  g = the origin in origin-grid steps
// (org = p0 + g*step, step = the grid step or 1 A), unwrapped.
void system_origin(const double *x, int n, const smr_gridff *grid, bool surface, double org[3], int &iux, int &iuy,
                   int g[3]) {
    double c[3] = {0, 0, 0};
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) c[k] += x[3 * i + k];
    for (int k = 0; k < 3; k++) c[k] /= n;
    if (grid && surface) {
        const smr_grid_desc &d = grid->desc;
        long long jx = std::llround((c[0] - d.p0[0]) / d.step[0]);
        long long jy = std::llround((c[1] - d.p0[1]) / d.step[1]);
        org[0] = d.p0[0] + jx * d.step[0];
        org[1] = d.p0[1] + jy * d.step[1];
        org[2] = d.p0[2];
        iux = (int)(((jx % d.nx) + d.nx) % d.nx);
        iuy = (int)(((jy % d.ny) + d.ny) % d.ny);
        g[0] = (int)jx;
        g[1] = (int)jy;
        g[2] = 0;
    } else {
        for (int k = 0; k < 3; k++) {
            org[k] = std::round(c[k]);
            g[k] = (int)org[k];
        }
        iux = iuy = 0;
    }
}

// workgroups of a shape that the device holds at once
static long long shape_capacity(const smr_ctx *c, ClassPlan p, const LaunchShape &L, bool have_grid) {
    p.count = 1 << 30;
    const LaunchShape m = make_shape(c, p, L.block, L.apt, L.fint_lds, L.pair_mode, L.stencil_cache, have_grid);
    return m.block ? m.grid : 0;
}

// launch shapes of a batch's classes (after a change of the tuning knobs),
// then the launches: adjacent classes (contiguous slices of the work list)
// whose shapes agree share one persistent launch when the merged launch
// keeps the resident workgroups of each (so the larger class's long
// relaxations start first and the smaller class fills in, instead of two
// launches each ending in its own tail)
void refresh_shapes(smr_batch *b) {
    const bool g = b->grid != nullptr;
    for (auto &p : b->class_plans)
        p.shape = class_shape(b->ctx, p, g, b->tune_block, b->tune_cache, b->tune_pairs);
    b->plans.clear();
    for (const auto &p : b->class_plans) {
        if (!b->plans.empty() && p.shape.block) {
            ClassPlan &q = b->plans.back();
            const LaunchShape &a = q.shape, &c = p.shape;
            if (a.block == c.block && a.apt == c.apt && a.fint_lds == c.fint_lds && a.pair_mode == c.pair_mode &&
                a.stencil_cache == c.stencil_cache && q.first + q.count == p.first) {
                ClassPlan m = q;
                m.count += p.count;
                m.max_atoms = std::max(q.max_atoms, p.max_atoms);
                m.nf_max = std::max(q.nf_max, p.nf_max);
                m.na_max = std::max(q.na_max, p.na_max);
                m.nd_max = std::max(q.nd_max, p.nd_max);
                m.ni_max = std::max(q.ni_max, p.ni_max);
                m.shape = make_shape(b->ctx, m, a.block, a.apt, a.fint_lds, a.pair_mode, a.stencil_cache, g);
                if (m.shape.block &&
                    shape_capacity(b->ctx, m, m.shape, g) >=
                        std::max(shape_capacity(b->ctx, q, a, g), shape_capacity(b->ctx, p, c, g))) {
                    q = m;
                    continue;
                }
            }
        }
        b->plans.push_back(p);
    }
    b->ls.drop_graph();
}

}  // namespace smr

namespace smr {

// Coupled layer set-up: per-atom system ids and REQ rows, the cell grid over
// the initial positions (cells >= the cutoff, 1 A margin; atoms that leave
// the box are binned into its edge cells, which keeps the 27-cell search
// exact), cell-list buffers and the shared FIRE state.
int layer_setup(smr_batch *b, const double *apos) {
    const double rc = b->opts.inter_cutoff;
    const int na = b->total_atoms, ns = b->nsys;
    std::vector<int> asys(na);
    std::vector<float4> areq(na);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    const double p0[3] = {b->grid ? b->grid->desc.p0[0] : 0.0, b->grid ? b->grid->desc.p0[1] : 0.0,
                          b->grid ? b->grid->desc.p0[2] : 0.0};
    for (int s = 0; s < ns; s++) {
        const TopoDesc &td = b->hp.topo[b->hp.sys[s].topo];
        for (int i = 0; i < td.natoms; i++) {
            const int a = b->atom_off[s] + i;
            asys[a] = s;
            areq[a] = b->hp.req[td.atom_off + i];
            for (int k = 0; k < 3; k++) {
                // the origin-grid frame of the kernels: global position - p0
                const double x = apos[3 * (size_t)a + k] - p0[k];
                lo[k] = std::min(lo[k], x);
                hi[k] = std::max(hi[k], x);
            }
        }
    }
    // atoms that leave the box are binned into its edge cells (exact for the
    // 27-cell search), so a small margin suffices
    const double margin = 1.0;
    double h = rc * (1.0 + 1e-4) + 1e-3;
    // periodic axes: the box [0, L) of the origin-grid frame in whole cells
    // of at least the cutoff, at least 3 of them (one image per pair)
    const double Lb[2] = {b->opts.pbc_lx, b->opts.pbc_ly};
    const double ostep[2] = {b->tab.ostep64_x, b->tab.ostep64_y};
    int nL[2] = {0, 0}, np_[2] = {0, 0};
    for (int k = 0; k < 2; k++) {
        if (Lb[k] <= 0.0) continue;
        const char *ax = k ? "pbc_ly" : "pbc_lx";
        np_[k] = (int)std::floor(Lb[k] / h);
        if (np_[k] < 3)
            return fail(SMR_ERR_INVALID, std::string("smr_batch_create: ") + ax + " = " + std::to_string(Lb[k]) +
                                             " A holds fewer than 3 cells of the cutoff " + std::to_string(rc) + " A");
        const double q = Lb[k] / ostep[k];
        nL[k] = (int)std::lround(q);
        if (std::fabs(q - nL[k]) > 1e-6 * std::max(1.0, q))
            return fail(SMR_ERR_INVALID, std::string("smr_batch_create: ") + ax + " = " + std::to_string(Lb[k]) +
                                             " A is not a whole number of origin-grid steps (" +
                                             std::to_string(ostep[k]) + " A)");
        if (b->grid) {
            const double period = (k ? b->grid->desc.ny * b->grid->desc.step[1] : b->grid->desc.nx * b->grid->desc.step[0]);
            const double m = Lb[k] / period;
            if (std::fabs(m - std::lround(m)) > 1e-6 * std::max(1.0, m))
                return fail(SMR_ERR_INVALID, std::string("smr_batch_create: ") + ax + " = " + std::to_string(Lb[k]) +
                                                 " A is not a whole number of GridFF periods (" +
                                                 std::to_string(period) + " A)");
        }
    }
    int n[3];
    for (;;) {
        long long tot = 1;
        for (int k = 0; k < 3; k++) {
            n[k] = (k < 2 && np_[k]) ? np_[k] : std::max(1, (int)std::ceil((hi[k] - lo[k] + 2.0 * margin) / h));
            tot *= n[k];
        }
        if (tot <= (1 << 20)) break;
        h *= 1.5;   // a very sparse layer: fewer, larger cells
        for (int k = 0; k < 2; k++)
            if (np_[k]) np_[k] = std::max(3, (int)std::floor(Lb[k] / h));
    }
    CellGrid &C = b->cells;
    memset(&C, 0, sizeof(C));
    C.ncx = n[0]; C.ncy = n[1]; C.ncz = n[2];
    C.nox = n[0] > 1 ? 3 : 1;
    C.noy = n[1] > 1 ? 3 : 1;
    C.noz = n[2] > 1 ? 3 : 1;
    C.lox = np_[0] ? 0.f : (float)(lo[0] - margin);
    C.loy = np_[1] ? 0.f : (float)(lo[1] - margin);
    C.loz = (float)(lo[2] - margin);
    C.inv_h = (float)(1.0 / h);
    C.inv_hx = np_[0] ? (float)(n[0] / Lb[0]) : C.inv_h;
    C.inv_hy = np_[1] ? (float)(n[1] / Lb[1]) : C.inv_h;
    C.pbc_x = np_[0] ? 1 : 0;
    C.pbc_y = np_[1] ? 1 : 0;
    C.Lx = (float)Lb[0];
    C.Ly = (float)Lb[1];
    C.inv_Lx = np_[0] ? (float)(1.0 / Lb[0]) : 0.f;
    C.inv_Ly = np_[1] ? (float)(1.0 / Lb[1]) : 0.f;
    C.nLx = nL[0];
    C.nLy = nL[1];
    C.rc2 = (float)(rc * rc);
    C.rc2_64 = rc * rc;
    const double w = b->opts.inter_switch;
    C.w64 = w;
    C.ron64 = rc - w;
    C.ron = (float)(rc - w);
    C.ron2 = w > 0.0 ? (float)((rc - w) * (rc - w)) : C.rc2;
    C.inv_w = w > 0.0 ? (float)(1.0 / w) : 0.f;
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    hipStream_t st = b->ctx->stream;
    hipError_t e = b->atom_sys.upload(asys.data(), sizeof(int) * na, st);
    if (e == hipSuccess) e = b->req_atom.upload(areq.data(), sizeof(float4) * na, st);
    if (e == hipSuccess) e = b->cell_count.alloc(sizeof(int) * ncell);
    if (e == hipSuccess) e = hipMemsetAsync(b->cell_count.ptr, 0, sizeof(int) * ncell, st);
    if (e == hipSuccess) e = b->cell_start.alloc(sizeof(int) * (ncell + 1));
    if (e == hipSuccess) e = b->cell_sidx.alloc(sizeof(int) * na);
    if (e == hipSuccess) e = b->cell_spos.alloc(sizeof(float4) * na);
    if (e == hipSuccess) e = b->cell_sreq.alloc(sizeof(float4) * na);
    if (e == hipSuccess) e = b->atom_cell.alloc(sizeof(int) * na);
    if (e == hipSuccess) e = b->atom_wrap.alloc(sizeof(int2) * na);
    if (e == hipSuccess) e = hipMemsetAsync(b->atom_wrap.ptr, 0, sizeof(int2) * na, st);
    if (e == hipSuccess) e = b->cell_swrap.alloc(sizeof(int2) * na);
    if (e == hipSuccess) e = hipMemsetAsync(b->cell_swrap.ptr, 0, sizeof(int2) * na, st);
    if (e == hipSuccess) e = b->cell_part.alloc(sizeof(float4) * 27 * (size_t)na);
    if (e == hipSuccess) e = b->slot_of.alloc(sizeof(int) * na);
    if (e == hipSuccess) e = hipMemsetAsync(b->slot_of.ptr, 0, sizeof(int) * na, st);
    if (e == hipSuccess) e = b->partials.alloc(sizeof(double4) * ns);
    if (e == hipSuccess) e = b->layer_state.alloc(sizeof(LayerState));
    if (e == hipSuccess) e = hipMemsetAsync(b->layer_state.ptr, 0, sizeof(LayerState), st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "smr_batch_create: layer buffers");
    C.spos = (const float4 *)b->cell_spos.ptr;
    C.sreq = (const float4 *)b->cell_sreq.ptr;
    C.start = (const int *)b->cell_start.ptr;
    C.part = (const float4 *)b->cell_part.ptr;
    C.slot_of = (const int *)b->slot_of.ptr;
    C.swrap = (const int2 *)b->cell_swrap.ptr;
    C.natoms = na;
    b->layer = true;
    return SMR_OK;
}

LayerArgs layer_args(smr_batch *b) {
    LayerArgs L;
    memset(&L, 0, sizeof(L));
    L.tab = b->tab;
    L.pos = (float4 *)b->pos.ptr;
    L.vel = (float4 *)b->vel.ptr;
    L.force = (const float4 *)b->force.ptr;
    L.atom_sys = (const int *)b->atom_sys.ptr;
    L.req_atom = (const float4 *)b->req_atom.ptr;
    L.natoms = b->total_atoms;
    L.nsys = b->nsys;
    L.count = (int *)b->cell_count.ptr;
    L.start = (int *)b->cell_start.ptr;
    L.sidx = (int *)b->cell_sidx.ptr;
    L.spos = (float4 *)b->cell_spos.ptr;
    L.sreq = (float4 *)b->cell_sreq.ptr;
    L.atom_cell = (int *)b->atom_cell.ptr;
    L.atom_wrap = (int2 *)b->atom_wrap.ptr;
    L.swrap = (int2 *)b->cell_swrap.ptr;
    L.part = (float4 *)b->cell_part.ptr;
    L.slot_of = (int *)b->slot_of.ptr;
    L.cells = b->cells;
    L.layer = (LayerState *)b->layer_state.ptr;
    L.work_counters = (int *)b->ls.counters.ptr;
    L.nwork_counters = NCLASS;
    return L;
}

// (re)build the cell list from the current positions, no move
int layer_cells(smr_batch *b) {
    hipError_t e = launch_layer_cells(layer_args(b), false, b->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "layer cell list");
    return SMR_OK;
}

}  // namespace smr

extern "C" int smr_batch_create(smr_ctx *c, const smr_gridff *grid, int ntopo, const smr_topology *topos, int nsys,
                                const int *sys_topo, const double *apos, const smr_ff_options *opts, smr_batch **out) {
    if (!c || !out || !topos || !sys_topo || !apos || !opts)
        return fail(SMR_ERR_INVALID, "smr_batch_create: NULL argument");
    *out = nullptr;
    if (ntopo <= 0 || nsys <= 0) return fail(SMR_ERR_INVALID, "smr_batch_create: ntopo and nsys must be > 0");
    if (opts->surface && !grid) return fail(SMR_ERR_INVALID, "smr_batch_create: surface requested but grid == NULL");
    if (!(opts->rdamp >= 0.0)) return fail(SMR_ERR_INVALID, "smr_batch_create: rdamp must be >= 0");
    if (!(opts->inter_cutoff >= 0.0) || opts->inter_cutoff > 1000.0)
        return fail(SMR_ERR_INVALID, "smr_batch_create: inter_cutoff must be in [0, 1000] A");
    if (opts->inter_cutoff > 0.0 && !opts->nonbonded)
        return fail(SMR_ERR_INVALID, "smr_batch_create: a coupled layer (inter_cutoff > 0) needs nonbonded = 1");
    if (!(opts->inter_switch >= 0.0) || (opts->inter_switch > 0.0 && opts->inter_switch >= opts->inter_cutoff))
        return fail(SMR_ERR_INVALID, "smr_batch_create: inter_switch must be in [0, inter_cutoff)");
    if (!(opts->pbc_lx >= 0.0) || !(opts->pbc_ly >= 0.0) || opts->pbc_lx > 1e6 || opts->pbc_ly > 1e6)
        return fail(SMR_ERR_INVALID, "smr_batch_create: pbc_lx, pbc_ly must be in [0, 1e6] A");
    if ((opts->pbc_lx > 0.0 || opts->pbc_ly > 0.0) && !(opts->inter_cutoff > 0.0))
        return fail(SMR_ERR_INVALID, "smr_batch_create: a periodic box needs a coupled layer (inter_cutoff > 0)");
    HIPCHK(hipSetDevice(c->device));
    auto *b = new smr_batch();
    b->ctx = c;
    ctx_retain(c);
    b->grid = opts->surface ? grid : nullptr;
    b->opts = *opts;
    const double alpha = grid ? grid->desc.alpha_morse : 0.0;
    for (int t = 0; t < ntopo; t++) {
        TopoDesc td;
        int rc = pack_topology(topos[t], t, alpha, b->grid != nullptr, b->hp, td);
        if (rc) { delete b; ctx_release(c); return rc; }
        b->hp.topo.push_back(td);
        b->max_atoms = std::max(b->max_atoms, td.natoms);
        b->nf_max = std::max(b->nf_max, td.nf);
        b->na_max = std::max(b->na_max, td.nangles);
        b->nd_max = std::max(b->nd_max, td.ndih);
        b->ni_max = std::max(b->ni_max, td.ninv);
    }
    b->nsys = nsys;
    b->atom_off.assign(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) {
        int t = sys_topo[s];
        if (t < 0 || t >= ntopo) {
            delete b;
            ctx_release(c);
            return fail(SMR_ERR_INVALID, "smr_batch_create: sys_topo[" + std::to_string(s) + "] = " + std::to_string(t) +
                                             " not in [0," + std::to_string(ntopo) + ")");
        }
        b->atom_off[s + 1] = b->atom_off[s] + b->hp.topo[t].natoms;
    }
    b->total_atoms = b->atom_off[nsys];
    b->origin.assign(3 * (size_t)nsys, 0.0);
    for (int s = 0; s < nsys; s++) {
        SysDesc sd;
        sd.topo = sys_topo[s];
        sd.atom_base = b->atom_off[s];
        int g[3];
        system_origin(apos + 3 * (size_t)b->atom_off[s], b->hp.topo[sd.topo].natoms, b->grid, b->grid != nullptr,
                      &b->origin[3 * s], sd.iux, sd.iuy, g);
        sd.gx = g[0];
        sd.gy = g[1];
        sd.gz = g[2];
        sd.pad = 0;
        b->hp.sys.push_back(sd);
    }
    hipError_t e = upload_pools(b->hp, b->dp, c->stream);
    if (e == hipSuccess) e = b->pos.alloc(sizeof(float4) * b->total_atoms);
    if (e == hipSuccess) e = b->vel.alloc(sizeof(float4) * b->total_atoms);
    if (e == hipSuccess) e = b->force.alloc(sizeof(float4) * b->total_atoms);
    if (e == hipSuccess) e = b->state.alloc(sizeof(SysState) * nsys);
    if (e == hipSuccess) e = hipMemsetAsync(b->state.ptr, 0, sizeof(SysState) * nsys, c->stream);
    if (e == hipSuccess) e = b->work.alloc(sizeof(int) * nsys);
    if (e == hipSuccess && b->grid) e = b->stencil.alloc(sizeof(float4) * 16 * (size_t)b->total_atoms);
    if (e == hipSuccess) e = b->stencil_key.alloc(sizeof(int) * (size_t)b->total_atoms);
    if (e == hipSuccess) e = hipMemsetAsync(b->stencil_key.ptr, 0xff, sizeof(int) * (size_t)b->total_atoms, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->vel.ptr, 0, sizeof(float4) * b->total_atoms, c->stream);
    if (e != hipSuccess) { delete b; ctx_release(c); return hip_fail(e, "smr_batch_create: device allocation"); }
    fill_tables(b->dp, b->grid, b->opts, b->tab);
    int rc = smr_batch_set_positions(b, apos);
    if (rc) { delete b; ctx_release(c); return rc; }
    e = b->pos0.alloc(sizeof(float4) * b->total_atoms);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->pos0.ptr, b->pos.ptr, sizeof(float4) * b->total_atoms, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) { delete b; ctx_release(c); return hip_fail(e, "smr_batch_create: initial positions"); }
    std::vector<SysSize> sizes(nsys);
    for (int s = 0; s < nsys; s++) {
        const TopoDesc &td = b->hp.topo[sys_topo[s]];
        sizes[s] = SysSize{s, td.natoms, td.nf, td.nangles, td.ndih, td.ninv};
    }
    std::vector<int> wl;
    build_plans(sizes, b->class_plans, wl);
    refresh_shapes(b);
    b->work_host = wl;
    e = hipMemcpyAsync(b->work.ptr, wl.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = b->work_base.alloc(sizeof(int) * nsys);
    if (e == hipSuccess) e = hipMemcpyAsync(b->work_base.ptr, wl.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = b->hstate.alloc(sizeof(SysState) * nsys);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { delete b; ctx_release(c); return hip_fail(e, "smr_batch_create: upload"); }
    if (opts->inter_cutoff > 0.0) {
        rc = layer_setup(b, apos);
        if (rc) { delete b; ctx_release(c); return rc; }
    }
    *out = b;
    return SMR_OK;
}

smr_batch::~smr_batch() {
    for (hipEvent_t &e : ev)
        if (e) (void)hipEventDestroy(e);
}

extern "C" int smr_batch_destroy(smr_batch *b) {
    if (!b) return SMR_OK;
    smr_ctx *c = b->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete b;
    ctx_release(c);
    return SMR_OK;
}

extern "C" int smr_batch_natoms(const smr_batch *b, int *total) {
    if (!b || !total) return fail(SMR_ERR_INVALID, "smr_batch_natoms: NULL argument");
    *total = b->total_atoms;
    return SMR_OK;
}

extern "C" int smr_batch_atom_offsets(const smr_batch *b, int *offsets) {
    if (!b || !offsets) return fail(SMR_ERR_INVALID, "smr_batch_atom_offsets: NULL argument");
    std::copy(b->atom_off.begin(), b->atom_off.end(), offsets);
    return SMR_OK;
}

extern "C" int smr_batch_set_positions(smr_batch *b, const double *apos) {
    if (!b || !apos) return fail(SMR_ERR_INVALID, "smr_batch_set_positions: NULL argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    std::vector<float4> h(b->total_atoms);
    for (int s = 0; s < b->nsys; s++) {
        const double *o = &b->origin[3 * s];
        for (int i = b->atom_off[s]; i < b->atom_off[s + 1]; i++) {
            double x = apos[3 * (size_t)i] - o[0], y = apos[3 * (size_t)i + 1] - o[1], z = apos[3 * (size_t)i + 2] - o[2];
            if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
                return fail(SMR_ERR_INVALID, "smr_batch_set_positions: atom " + std::to_string(i) + " is not finite");
            h[i] = make_float4((float)x, (float)y, (float)z, 0.f);
        }
    }
    HIPCHK(hipMemcpyAsync(b->pos.ptr, h.data(), sizeof(float4) * h.size(), hipMemcpyHostToDevice, b->ctx->stream));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    return SMR_OK;
}

extern "C" int smr_batch_get_positions(const smr_batch *b, double *apos) {
    if (!b || !apos) return fail(SMR_ERR_INVALID, "smr_batch_get_positions: NULL argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    std::vector<float4> h(b->total_atoms);
    HIPCHK(hipMemcpyAsync(h.data(), b->pos.ptr, sizeof(float4) * h.size(), hipMemcpyDeviceToHost, b->ctx->stream));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    for (int s = 0; s < b->nsys; s++) {
        const double *o = &b->origin[3 * s];
        for (int i = b->atom_off[s]; i < b->atom_off[s + 1]; i++) {
            apos[3 * (size_t)i] = (double)h[i].x + o[0];
            apos[3 * (size_t)i + 1] = (double)h[i].y + o[1];
            apos[3 * (size_t)i + 2] = (double)h[i].z + o[2];
        }
    }
    return SMR_OK;
}

// ============================================================ kernels drivers
namespace smr {

KernelArgs fused_args(smr_batch *b, int mode, const FireConst &fc, int steps_per_launch) {
    KernelArgs A;
    memset(&A, 0, sizeof(A));
    A.tab = b->tab;
    A.fire = fc;
    A.pos = (float4 *)b->pos.ptr;
    A.vel = (float4 *)b->vel.ptr;
    A.force = (float4 *)b->force.ptr;
    A.state = (SysState *)b->state.ptr;
    A.work_list = (const int *)b->work.ptr;
    A.stencil = (float4 *)b->stencil.ptr;
    A.stencil_key = (int *)b->stencil_key.ptr;
    if (b->layer) A.cells = b->cells;
    A.steps_per_launch = steps_per_launch;
    A.mode = mode;
    return A;
}

int run_fused(smr_batch *b, int mode, const FireConst &fc, int steps_per_launch, bool reset_state,
              double dt0, double alpha0, float *kernel_ms) {
    smr_ctx *c = b->ctx;
    if (reset_state) HIPCHK(launch_reset_state((SysState *)b->state.ptr, b->nsys, dt0, alpha0, c->stream));
    return launch_plans(c->ss, b->ls, b->plans, fused_args(b, mode, fc, steps_per_launch), kernel_ms);
}

// FP64 energies of every system at its current positions (into state[].E)
// Longest-expected-first order of the work list: inside each size class the
// larger systems come first (as built), and among systems of one size those
// whose starting positions have the highest FP64 energy per atom (the
// furthest from a minimum, which need the most FIRE steps), so that long
// relaxations do not start in the tail of the persistent launch.  The order
// affects only the schedule: every system's trajectory is its own.
// (only where a size class has more systems than resident workgroups:
// otherwise every system starts at once anyway)
bool reorder_pays(const smr_batch *b) {
    if (!b->energy_order || b->layer || b->nsys < 2) return false;
    for (const auto &p : b->plans)
        if (p.count > p.shape.grid) return true;
    return false;
}

int order_by_energy(smr_batch *b) {
    if (!reorder_pays(b)) return SMR_OK;
    hipStream_t st = b->ctx->stream;
    // asynchronous on the batch's stream: FP64 energies of the current
    // positions, then the ranked work list (order_kernels.hip)
    hipError_t e = launch_energy64(b->tab, (const float4 *)b->pos.ptr, (SysState *)b->state.ptr,
                                   (const int *)b->work_base.ptr, b->nsys, st);
    RankSlices rs;
    rs.n = (int)b->plans.size();
    for (int q = 0; q < rs.n; q++) rs.first[q] = b->plans[q].first;
    rs.first[rs.n] = b->nsys;
    if (e == hipSuccess)
        e = rank_work_list(b->tab, (const SysState *)b->state.ptr, (const int *)b->work_base.ptr, b->nsys, rs, b->rank,
                           (int *)b->work.ptr, st);
    if (e != hipSuccess) return hip_fail(e, "order_by_energy");
    return SMR_OK;
}

int eval_energies64(smr_batch *b) {
    hipStream_t st = b->ctx->stream;
    hipError_t e = launch_energy64(b->tab, (const float4 *)b->pos.ptr, (SysState *)b->state.ptr,
                                   (const int *)b->work.ptr, b->nsys, st);
    if (e == hipSuccess && b->layer)
        e = launch_energy64_inter(layer_args(b), (SysState *)b->state.ptr, (const int *)b->work.ptr, b->nsys, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "k_energy64");
    return SMR_OK;
}

void results_from_states(const SysState *h, int nsys, smr_relax_result *res) {
    for (int s = 0; s < nsys; s++) {
        res[s].steps = h[s].steps;
        res[s].status = h[s].status;
        res[s].fmax = h[s].fmax;
        double E = 0.0;
        for (int q = 0; q < 6; q++) { res[s].terms[q] = h[s].E[q]; E += h[s].E[q]; }
        res[s].energy = E;
    }
}

int read_results(smr_batch *b, smr_relax_result *res) {
    std::vector<SysState> h(b->nsys);
    HIPCHK(hipMemcpy(h.data(), b->state.ptr, sizeof(SysState) * h.size(), hipMemcpyDeviceToHost));
    results_from_states(h.data(), b->nsys, res);
    return SMR_OK;
}

int check_fire(const smr_fire_params *f, const char *who) {
    if (!f) return fail(SMR_ERR_INVALID, std::string(who) + ": fire == NULL");
    if (!(f->dt0 > 0.0) || !(f->dt_max >= f->dt0) || !(f->f_dec > 0.0 && f->f_dec < 1.0) || !(f->f_inc >= 1.0) ||
        !(f->alpha0 >= 0.0 && f->alpha0 <= 1.0) || !(f->f_alpha > 0.0 && f->f_alpha <= 1.0) || f->n_min < 0)
        return fail(SMR_ERR_INVALID, std::string(who) + ": FIRE parameters out of range (dt0=" + std::to_string(f->dt0) +
                                         " dt_max=" + std::to_string(f->dt_max) + ")");
    return SMR_OK;
}

FireConst make_fire_const(const smr_fire_params *f, int max_steps, double fconv) {
    FireConst fc;
    memset(&fc, 0, sizeof(fc));
    if (f) {
        fc.dt0 = f->dt0; fc.dt_max = f->dt_max; fc.alpha0 = f->alpha0; fc.f_inc = f->f_inc;
        fc.f_dec = f->f_dec; fc.f_alpha = f->f_alpha; fc.flim = f->flim; fc.n_min = f->n_min;
    }
    fc.fconv = fconv;
    fc.max_steps = max_steps;
    return fc;
}

}  // namespace smr

namespace smr {

// FP64 energies of the final positions and the layer's shared step count,
// status and Fmax in every molecule's result.
int layer_results(smr_batch *b, const LayerState &hs, smr_relax_result *res) {
    int rc = eval_energies64(b);
    if (rc) return rc;
    rc = read_results(b, res);
    if (rc) return rc;
    const int status = hs.status == SMR_SYS_RUNNING ? SMR_SYS_MAXSTEPS : hs.status;
    for (int s = 0; s < b->nsys; s++) {
        res[s].steps = hs.steps;
        res[s].status = status;
        res[s].fmax = hs.fmax;
    }
    return SMR_OK;
}

// The coupled layer as one cooperative launch (KMODE_LAYER_P of k_relax,
// layer_persist.h), from the state layer_relax has set up.  *done = false
// when the layer does not take this path: switched off, molecules of
// another launch shape, more molecules than the chip holds at once, or a
// neighbourhood that overflows the staging area or the lists (the launch
// then stops without writing anything back, so the step launches start from
// the same state).  b->last_layer_fallback says which.
constexpr int kLpMaxStaged = 768;   // neighbour atoms staged per molecule
static_assert(kLpMaxStaged <= 1024 && kLpMaxInst <= 64, "a list entry packs a staged atom in 10 bits, its instance in 6");
constexpr int kLpMaxList = 256;     // list entries per atom at most (fewer when LDS is short)
constexpr int kLpMinList = 48;      // ... and at least
constexpr size_t kLdsPerCu = 160 * 1024;

int layer_relax_persistent(smr_batch *b, const std::vector<KernelArgs> &args, bool &done) {
    done = false;
    b->last_layer_persistent = 0;
    b->last_layer_fallback = 0;
    b->last_layer_rebuilds = 0;
    if (!b->layer_persistent) return SMR_OK;
    smr_ctx *c = b->ctx;
    hipStream_t st = c->stream;
    if (b->plans.size() != 1 || b->plans[0].shape.block != 64 || b->plans[0].shape.apt != 1 ||
        !b->plans[0].shape.fint_lds || b->plans[0].count != b->nsys) {
        b->last_layer_fallback = 6;
        return SMR_OK;
    }
    const ClassPlan &p = b->plans[0];
    const size_t off = (p.shape.lds_bytes + 15) & ~(size_t)15;
    // the lists take what LDS is left when ceil(nsys / CUs) molecules share a CU
    const int lds_atoms = args[0].lds_atoms;
    const int need_per_cu = (b->nsys + c->num_cu - 1) / c->num_cu;
    const size_t cap = std::min(c->max_lds, kLdsPerCu / (size_t)need_per_cu);
    const size_t fixed = off + lp_lds_bytes(kLpMaxStaged, lds_atoms, 0);
    const long long room = (long long)cap - (long long)fixed - 16;
    const int maxl = room > 0 ? (int)std::min<long long>(kLpMaxList, room / (2LL * lds_atoms)) : 0;
    if (maxl < kLpMinList) {
        b->last_layer_fallback = 5;
        return SMR_OK;
    }
    const size_t lds = off + lp_lds_bytes(kLpMaxStaged, lds_atoms, maxl);
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, relax_layer_persistent_ptr(), 64, lds);
    if (e != hipSuccess) return hip_fail(e, "occupancy of the one-launch layer kernel");
    if (lds > c->max_lds || (long long)per_cu * c->num_cu < b->nsys) {
        b->last_layer_fallback = 5;
        return SMR_OK;
    }
    const size_t na = b->total_atoms, ns = b->nsys;
    const size_t nbar = kLpBarWords;
    if (b->lp_xbuf.bytes < sizeof(float4) * 6 * na && (e = b->lp_xbuf.alloc(sizeof(float4) * 6 * na)) != hipSuccess)
        return hip_fail(e, "layer exchange buffer");
    if (b->lp_xrec.bytes < sizeof(double) * 2 * kLpRec * ns &&
        (e = b->lp_xrec.alloc(sizeof(double) * 2 * kLpRec * ns)) != hipSuccess)
        return hip_fail(e, "layer record buffer");
    if (b->lp_bar.bytes < sizeof(unsigned) * nbar && (e = b->lp_bar.alloc(sizeof(unsigned) * nbar)) != hipSuccess)
        return hip_fail(e, "layer barrier words");
    KernelArgs A = args[0];
    A.mode = KMODE_LAYER_P;
    A.lp.xbuf = (float4 *)b->lp_xbuf.ptr;
    A.lp.xrec = (double *)b->lp_xrec.ptr;
    A.lp.bar = (unsigned *)b->lp_bar.ptr;
    A.lp.maxl = maxl;
    A.lp.maxs = kLpMaxStaged;
    A.lp.lds_off = (int)off;
    A.lp.natoms = b->total_atoms;
    A.lp.skin = b->layer_skin;
    // barrier counters and the error word start at zero (a kernel, not a memset)
    if ((e = launch_zero_i32((int *)b->lp_bar.ptr, (int)nbar, st)) != hipSuccess) return hip_fail(e, "layer barrier reset");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = launch_relax_layer_persistent(A, b->nsys, lds, st);
    if (e == hipErrorCooperativeLaunchTooLarge) {
        (void)hipGetLastError();
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->last_layer_fallback = 5;
        return SMR_OK;
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&b->last_kernel_ms, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "one-launch layer relaxation");
    unsigned err = 0;
    HIPCHK(hipMemcpy(&err, (unsigned *)b->lp_bar.ptr + (kLpGroups + 1) * 32, sizeof(err), hipMemcpyDeviceToHost));
    if (err != 0u) {
        b->last_layer_fallback = (int)err;
        return SMR_OK;
    }
    LayerState hs;
    HIPCHK(hipMemcpy(&hs, b->layer_state.ptr, sizeof(hs), hipMemcpyDeviceToHost));
    b->last_layer_rebuilds = hs.rebuilds;
    b->last_layer_persistent = 1;
    done = true;
    return SMR_OK;
}

// FIRE on a coupled layer: one cooperative launch where the layer fits
// (layer_relax_persistent), else steps of [force launches (per size class)
// -> FIRE decision -> move + cell-list rebuild] captured S at a time into a
// HIP graph, which the host replays until the shared FIRE state leaves
// RUNNING.
int layer_relax(smr_batch *b, const smr_fire_params *fire, int max_steps, double fconv, smr_relax_result *res) {
    smr_ctx *c = b->ctx;
    hipStream_t st = c->stream;
    constexpr int S = 16;
    LayerState L0;
    memset(&L0, 0, sizeof(L0));
    L0.dt = fire->dt0;
    L0.alpha = fire->alpha0;
    L0.status = SMR_SYS_RUNNING;
    L0.nsys = b->nsys;
    std::vector<SysState> h(b->nsys);
    for (auto &x : h) {
        memset(&x, 0, sizeof(x));
        x.dt = fire->dt0;
        x.alpha = fire->alpha0;
        x.status = SMR_SYS_RUNNING;
    }
    HIPCHK(hipMemcpyAsync(b->layer_state.ptr, &L0, sizeof(L0), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->state.ptr, h.data(), sizeof(SysState) * h.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(b->vel.ptr, 0, sizeof(float4) * b->total_atoms, st));
    int rc = layer_cells(b);
    if (rc) return rc;
    KernelArgs A;
    memset(&A, 0, sizeof(A));
    A.tab = b->tab;
    A.fire = make_fire_const(fire, max_steps, fconv);
    A.pos = (float4 *)b->pos.ptr;
    A.vel = (float4 *)b->vel.ptr;
    A.force = (float4 *)b->force.ptr;
    A.state = (SysState *)b->state.ptr;
    A.work_list = (const int *)b->work.ptr;
    A.stencil = (float4 *)b->stencil.ptr;
    A.stencil_key = (int *)b->stencil_key.ptr;
    A.cells = b->cells;
    A.layer = (LayerState *)b->layer_state.ptr;
    A.partials = (double4 *)b->partials.ptr;
    A.steps_per_launch = 1;
    A.mode = KMODE_LAYER;
    std::vector<KernelArgs> args;
    rc = prepare_plans(b->ls, b->plans, A, args);
    if (rc) return rc;
    {
        KernelArgs P = A;
        P.steps_per_launch = max_steps;
        std::vector<KernelArgs> pargs;
        rc = prepare_plans(b->ls, b->plans, P, pargs);
        if (rc) return rc;
        bool one_launch = false;
        rc = layer_relax_persistent(b, pargs, one_launch);
        if (rc) return rc;
        if (one_launch) {
            LayerState hs;
            HIPCHK(hipMemcpy(&hs, b->layer_state.ptr, sizeof(hs), hipMemcpyDeviceToHost));
            // the cell list of the final positions, for the FP64 energies
            rc = layer_cells(b);
            if (rc) return rc;
            return layer_results(b, hs, res);
        }
    }
    LayerArgs LA = layer_args(b);
    LA.partials = A.partials;
    LA.fire = A.fire;
    // S steps of launches, captured once into a graph (or issued directly
    // when the batch's graphs are switched off)
    auto issue_steps = [&]() -> hipError_t {
        hipError_t ei = hipSuccess;
        for (int k = 0; k < S && ei == hipSuccess; k++) {
            ei = issue_plans(c->ss, b->ls, b->plans, args, false);
            if (ei == hipSuccess) ei = launch_layer_decide(LA, st);
            if (ei == hipSuccess) ei = launch_layer_cells(LA, true, st);
        }
        return ei;
    };
    const bool use_graph = b->ls.use_graph;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipError_t e = hipSuccess;
    if (use_graph) {
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const hipError_t ei = issue_steps();
        e = hipStreamEndCapture(st, &graph);
        if (ei != hipSuccess || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return hip_fail(ei != hipSuccess ? ei : e, "capture of the layer steps");
        }
        e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return hip_fail(e, "hipGraphInstantiate (layer steps)");
        }
    }
    // the first step's work counters (each step's cell build resets them for the next)
    if (e == hipSuccess) e = launch_zero_i32((int *)b->ls.counters.ptr, NCLASS, st);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    LayerState hs = L0;
    const int max_rep = max_steps / S + 2;
    for (int r = 0; r < max_rep && e == hipSuccess; r++) {
        e = use_graph ? hipGraphLaunch(gexec, st) : issue_steps();
        if (e == hipSuccess) e = hipMemcpyAsync(&hs, b->layer_state.ptr, sizeof(hs), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (hs.status != SMR_SYS_RUNNING) break;
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&b->last_kernel_ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "layer relaxation");
    return layer_results(b, hs, res);
}

}  // namespace smr

extern "C" int smr_batch_eval_forces(smr_batch *b, double *forces, double *energies) {
    if (!b || !forces || !energies) return fail(SMR_ERR_INVALID, "smr_batch_eval_forces: NULL argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    // keep the FIRE state: eval mode reads but does not need it; back it up
    std::vector<SysState> keep(b->nsys);
    HIPCHK(hipMemcpy(keep.data(), b->state.ptr, sizeof(SysState) * keep.size(), hipMemcpyDeviceToHost));
    FireConst fc = make_fire_const(nullptr, 1, 0.0);
    int rc = b->layer ? layer_cells(b) : SMR_OK;
    if (rc) return rc;
    rc = run_fused(b, KMODE_EVAL, fc, 1, false, 0, 0, nullptr);
    if (rc) return rc;
    rc = eval_energies64(b);
    if (rc) return rc;
    std::vector<float4> h(b->total_atoms);
    HIPCHK(hipMemcpy(h.data(), b->force.ptr, sizeof(float4) * h.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < b->total_atoms; i++) {
        forces[3 * (size_t)i] = h[i].x;
        forces[3 * (size_t)i + 1] = h[i].y;
        forces[3 * (size_t)i + 2] = h[i].z;
    }
    std::vector<SysState> hs(b->nsys);
    HIPCHK(hipMemcpy(hs.data(), b->state.ptr, sizeof(SysState) * hs.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < b->nsys; s++)
        for (int q = 0; q < 6; q++) energies[6 * (size_t)s + q] = hs[s].E[q];
    HIPCHK(hipMemcpy(b->state.ptr, keep.data(), sizeof(SysState) * keep.size(), hipMemcpyHostToDevice));
    return SMR_OK;
}

extern "C" int smr_relax_batched(smr_batch *b, const smr_fire_params *fire, int max_steps, double fconv,
                                 smr_relax_result *results) {
    if (!b || !results) return fail(SMR_ERR_INVALID, "smr_relax_batched: NULL argument");
    int rc = check_fire(fire, "smr_relax_batched");
    if (rc) return rc;
    if (max_steps <= 0 || !(fconv > 0.0)) return fail(SMR_ERR_INVALID, "smr_relax_batched: need max_steps > 0, fconv > 0");
    HIPCHK(hipSetDevice(b->ctx->device));
    if (b->layer) return layer_relax(b, fire, max_steps, fconv, results);
    // Everything up to the results' readback is issued without waiting:
    // the start order, the launches and the final FP64 energies run back to
    // back on the batch's stream.
    hipStream_t st = b->ctx->stream;
    HIPCHK(hipMemsetAsync(b->vel.ptr, 0, sizeof(float4) * b->total_atoms, st));
    HIPCHK(launch_reset_state((SysState *)b->state.ptr, b->nsys, fire->dt0, fire->alpha0, st));
    rc = order_by_energy(b);
    if (rc) return rc;
    for (hipEvent_t &e : b->ev)
        if (!e) HIPCHK(hipEventCreate(&e));
    FireConst fc = make_fire_const(fire, max_steps, fconv);
    const int rerank = b->rerank_steps < 0 ? max_steps / 5 : b->rerank_steps;
    const bool two = rerank >= 50 && rerank < max_steps && reorder_pays(b);
    b->ls_rerank.use_graph = b->ls.use_graph;
    if (!b->plans.empty()) {
        // two launches when re-ranking: every system takes up to `rerank`
        // steps (the state, velocities and stored stencils persist as in the
        // server's chunks), then the unfinished ones are ranked again by the
        // energy of their current positions and continue; trajectories are
        // unchanged
        rc = launch_plans_async(b->ctx->ss, b->ls, b->plans, fused_args(b, KMODE_FIRE, fc, two ? rerank : max_steps),
                                b->ev[0], b->ev[1]);
        if (rc) return rc;
        if (two) {
            rc = order_by_energy(b);
            if (rc) return rc;
            rc = launch_plans_async(b->ctx->ss, b->ls_rerank, b->plans, fused_args(b, KMODE_FIRE, fc, max_steps),
                                    b->ev[2], b->ev[3]);
            if (rc) return rc;
        }
    }
    HIPCHK(launch_energy64(b->tab, (const float4 *)b->pos.ptr, (SysState *)b->state.ptr, (const int *)b->work.ptr,
                           b->nsys, st));
    HIPCHK(hipMemcpyAsync(b->hstate.ptr, b->state.ptr, sizeof(SysState) * b->nsys, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float k1 = 0.f, k2 = 0.f;
    if (!b->plans.empty()) {
        HIPCHK(hipEventElapsedTime(&k1, b->ev[0], b->ev[1]));
        if (two) HIPCHK(hipEventElapsedTime(&k2, b->ev[2], b->ev[3]));
    }
    b->last_kernel_ms = k1 + k2;
    results_from_states((const SysState *)b->hstate.ptr, b->nsys, results);
    return SMR_OK;
}

extern "C" int smr_run_md(smr_batch *b, int niter, double dt, double fconv, double flim, double damping, int *steps) {
    if (!b || !steps) return fail(SMR_ERR_INVALID, "smr_run_md: NULL argument");
    if (niter <= 0 || !(dt > 0.0) || !(flim > 0.0)) return fail(SMR_ERR_INVALID, "smr_run_md: need niter > 0, dt > 0, flim > 0");
    if (b->layer) return fail(SMR_ERR_UNSUPPORTED, "smr_run_md: not available for a coupled layer (use smr_relax_batched)");
    HIPCHK(hipSetDevice(b->ctx->device));
    FireConst fc = make_fire_const(nullptr, niter, fconv);
    fc.md_dt = dt;
    fc.md_flim = flim;
    fc.md_cdamp = std::max(0.0, 1.0 - damping);
    int rc = run_fused(b, KMODE_MD, fc, niter, true, dt, 0.0, nullptr);
    if (rc) return rc;
    std::vector<SysState> h(b->nsys);
    HIPCHK(hipMemcpy(h.data(), b->state.ptr, sizeof(SysState) * h.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < b->nsys; s++) steps[s] = h[s].steps;
    return SMR_OK;
}

namespace smr {
// The plain method on system isys.  Its kernels' scratch (hneigh, force
// pieces, per-atom energies, the reduction output) comes from `ext` when
// given (no allocation on this path), else from the batch's own buffers,
// allocated on the first call.
int relax_plain_run(smr_batch *b, int isys, const smr_fire_params *fire, int max_steps, double fconv,
                    smr_relax_result *result, const PlainScratch *ext, const char *who) {
    if (!b || !result) return fail(SMR_ERR_INVALID, std::string(who) + ": NULL argument");
    int rc = check_fire(fire, who);
    if (rc) return rc;
    if (isys < 0 || isys >= b->nsys)
        return fail(SMR_ERR_INVALID, std::string(who) + ": isys " + std::to_string(isys) + " not in [0," +
                                         std::to_string(b->nsys) + ")");
    if (max_steps <= 0 || !(fconv > 0.0)) return fail(SMR_ERR_INVALID, std::string(who) + ": need max_steps > 0, fconv > 0");
    if (b->layer) return fail(SMR_ERR_UNSUPPORTED, std::string(who) + ": not available for a coupled layer");
    smr_ctx *c = b->ctx;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const SysDesc &sd = b->hp.sys[isys];
    const TopoDesc &td = b->hp.topo[sd.topo];
    const int N = td.natoms;
    PlainArgs P;
    if (ext) {
        if (N > ext->max_atoms || td.nf > ext->max_pieces)
            return fail(SMR_ERR_UNSUPPORTED, std::string(who) + ": system " + std::to_string(isys) + " has " +
                                                 std::to_string(N) + " atoms and " + std::to_string(td.nf) +
                                                 " force pieces; the fixed scratch holds at most " +
                                                 std::to_string(ext->max_atoms) + " atoms and " +
                                                 std::to_string(ext->max_pieces) + " pieces");
        P.hneigh = ext->hneigh;
        P.fint = ext->fint;
        P.eatom = ext->eatom;
        P.esurf = ext->esurf;
        P.epair = ext->epair;
        P.cvf = ext->cvf;
    } else {
        if (b->plain_hneigh.bytes < sizeof(float4) * 4 * (size_t)b->max_atoms) {
            hipError_t e = b->plain_hneigh.alloc(sizeof(float4) * 4 * (size_t)b->max_atoms);
            if (e == hipSuccess) e = b->plain_fint.alloc(sizeof(float4) * (size_t)std::max(1, b->nf_max));
            if (e == hipSuccess) e = b->plain_e.alloc((sizeof(float) * 2 + sizeof(double)) * (size_t)b->max_atoms);
            if (e == hipSuccess) e = b->plain_cvf.alloc(sizeof(double) * 16);
            if (e != hipSuccess) return hip_fail(e, "smr_relax_plain: scratch");
        }
        P.hneigh = (float4 *)b->plain_hneigh.ptr;
        P.fint = (float4 *)b->plain_fint.ptr;
        P.eatom = (float *)b->plain_e.ptr;
        P.esurf = P.eatom + b->max_atoms;
        P.epair = (double *)(P.esurf + b->max_atoms);
        P.cvf = (double *)b->plain_cvf.ptr;
    }
    P.tab = b->tab;
    P.topo = sd.topo;
    P.iux = sd.iux;
    P.iuy = sd.iuy;
    P.pos = (float4 *)b->pos.ptr + sd.atom_base;
    P.vel = (float4 *)b->vel.ptr + sd.atom_base;
    P.force = (float4 *)b->force.ptr + sd.atom_base;
    HIPCHK(hipMemsetAsync(P.vel, 0, sizeof(float4) * N, st));
    double dt = fire->dt0, alpha = fire->alpha0;
    int npos = 0, steps = max_steps, status = SMR_SYS_MAXSTEPS;
    double cvf[10] = {0};
    for (int itr = 0; itr < max_steps; itr++) {
        hipError_t e = launch_plain_step_forces(P, N, td.nangles, td.ndih, td.ninv, st);
        if (e == hipSuccess) e = launch_plain_reduce(P, dt, fire->flim, st);
        if (e == hipSuccess) e = hipMemcpyAsync(cvf, P.cvf, sizeof(double) * 10, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "smr_relax_plain: step");
        const double vf = cvf[0], vv = cvf[1], ff = cvf[2];
        const double dt_old = dt;
        float a = 1.f, bb = 0.f;
        int reset = 0;
        if (!(std::isfinite(vf) && std::isfinite(vv) && std::isfinite(ff))) {
            steps = itr + 1;
            status = SMR_SYS_NONFINITE;
            break;
        }
        if (vf > 0.0 && ff > 1e-30 && vv > 1e-30) {
            a = (float)(1.0 - alpha);
            bb = (float)(alpha * std::sqrt(vv) / std::sqrt(ff));
            npos++;
            if (npos > fire->n_min) {
                dt = std::min(dt * fire->f_inc, fire->dt_max);
                alpha *= fire->f_alpha;
            }
        } else {
            npos = 0;
            dt *= fire->f_dec;
            alpha = fire->alpha0;
            reset = 1;
        }
        e = launch_plain_apply(P, N, a, bb, reset, (float)dt_old, st);
        if (e != hipSuccess) return hip_fail(e, "smr_relax_plain: apply");
        if (std::sqrt(cvf[3]) < fconv) {
            steps = itr + 1;
            status = SMR_SYS_CONVERGED;
            break;
        }
    }
    // energies of the final positions, FP64 (as for the batched path); the
    // system index travels in the reduction scratch
    HIPCHK(hipMemcpyAsync(P.cvf + 12, &isys, sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(launch_energy64(b->tab, (const float4 *)b->pos.ptr, (SysState *)b->state.ptr, (const int *)(P.cvf + 12), 1, st));
    SysState hs;
    HIPCHK(hipMemcpyAsync(&hs, (SysState *)b->state.ptr + isys, sizeof(SysState), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    result->steps = steps;
    result->status = status;
    result->fmax = std::sqrt(cvf[3]);
    double E = 0.0;
    for (int q = 0; q < 6; q++) { result->terms[q] = hs.E[q]; E += hs.E[q]; }
    result->energy = E;
    return SMR_OK;
}
}  // namespace smr

extern "C" int smr_relax_plain(smr_batch *b, int isys, const smr_fire_params *fire, int max_steps, double fconv,
                               smr_relax_result *result) {
    return relax_plain_run(b, isys, fire, max_steps, fconv, result, nullptr, "smr_relax_plain");
}

extern "C" int smr_batch_last_launch(const smr_batch *b, int *shape4, float *kernel_ms) {
    if (!b || !shape4 || !kernel_ms) return fail(SMR_ERR_INVALID, "smr_batch_last_launch: NULL argument");
    // the class with the most systems
    const ClassPlan *p = nullptr;
    for (const auto &q : b->plans)
        if (!p || q.count > p->count) p = &q;
    const LaunchShape L = p ? p->shape : LaunchShape();
    shape4[0] = L.block;
    shape4[1] = L.apt;
    shape4[2] = L.grid;
    shape4[3] = (int)L.lds_bytes | (L.stencil_cache ? (1 << 30) : 0);
    *kernel_ms = b->last_kernel_ms;
    return SMR_OK;
}

extern "C" int smr_batch_plan(const smr_batch *b, int *nclass, int *info) {
    if (!b || !nclass || !info) return fail(SMR_ERR_INVALID, "smr_batch_plan: NULL argument");
    *nclass = (int)b->plans.size();
    for (size_t q = 0; q < b->plans.size(); q++) {
        const ClassPlan &p = b->plans[q];
        int *o = info + 8 * q;
        o[0] = p.cls;
        o[1] = p.count;
        o[2] = p.max_atoms;
        o[3] = p.shape.block;
        o[4] = p.shape.apt;
        o[5] = p.shape.grid;
        o[6] = (int)p.shape.lds_bytes;
        o[7] = (p.shape.fint_lds ? 1 : 0) | (p.shape.pair_mode ? 2 : 0) | (p.shape.stencil_cache ? 4 : 0);
    }
    return SMR_OK;
}

// Autotuner: for every size class of the batch, time each launch shape that
// holds the class (threads per workgroup x atoms per thread at the class's
// capacity and one step up, pair sweep or ordered pairs, force pieces in LDS
// or in global scratch) on a FIRE run of trial_steps steps from the batch's
// current state, and keep the fastest in the context.  Positions, velocities
// and FIRE state are restored afterwards.
extern "C" int smr_batch_autotune(smr_batch *b, const smr_fire_params *fire, int trial_steps) {
    if (!b) return fail(SMR_ERR_INVALID, "smr_batch_autotune: NULL argument");
    int rc = check_fire(fire, "smr_batch_autotune");
    if (rc) return rc;
    if (trial_steps <= 0 || trial_steps > 100000)
        return fail(SMR_ERR_INVALID, "smr_batch_autotune: trial_steps must be in [1, 100000]");
    smr_ctx *c = b->ctx;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const size_t apos_b = sizeof(float4) * b->total_atoms, st_b = sizeof(SysState) * b->nsys;
    DevBuf kpos, kvel, kst;
    HIPCHK(kpos.alloc(apos_b));
    HIPCHK(kvel.alloc(apos_b));
    HIPCHK(kst.alloc(st_b));
    HIPCHK(hipMemcpyAsync(kpos.ptr, b->pos.ptr, apos_b, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(kvel.ptr, b->vel.ptr, apos_b, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(kst.ptr, b->state.ptr, st_b, hipMemcpyDeviceToDevice, st));
    std::vector<SysState> h0(b->nsys);
    for (auto &x : h0) {
        memset(&x, 0, sizeof(x));
        x.dt = fire->dt0;
        x.alpha = fire->alpha0;
        x.status = SMR_SYS_RUNNING;
    }
    // fconv = 0: nobody converges inside the trial
    FireConst fc = make_fire_const(fire, 1 << 30, 0.0);
    KernelArgs A;
    memset(&A, 0, sizeof(A));
    A.tab = b->tab;
    A.fire = fc;
    A.pos = (float4 *)b->pos.ptr;
    A.vel = (float4 *)b->vel.ptr;
    A.force = (float4 *)b->force.ptr;
    A.state = (SysState *)b->state.ptr;
    A.work_list = (const int *)b->work.ptr;
    A.stencil = (float4 *)b->stencil.ptr;
    A.steps_per_launch = trial_steps;
    A.mode = KMODE_FIRE;
    static const int shapes[][2] = {{64, 1}, {64, 2}, {64, 4}, {128, 1}, {128, 2}, {128, 4}, {256, 1}, {256, 2}, {256, 4},
                                    {512, 1}, {512, 2}};
    LaunchSet ls;
    ls.use_graph = false;
    auto trial = [&](const ClassPlan &p, float &ms) -> int {
        HIPCHK(hipMemcpyAsync(b->pos.ptr, kpos.ptr, apos_b, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemsetAsync(b->vel.ptr, 0, apos_b, st));
        HIPCHK(hipMemcpyAsync(b->state.ptr, h0.data(), st_b, hipMemcpyHostToDevice, st));
        std::vector<ClassPlan> one(1, p);
        return launch_plans(c->ss, ls, one, A, &ms);
    };
    int rc2 = SMR_OK;
    for (auto &p : b->class_plans) {
        const int cap0 = kClassMaxAtoms[p.cls];
        std::vector<LaunchShape> cand;
        for (auto &sh : shapes) {
            const int cap = sh[0] * sh[1];
            if (cap < p.max_atoms || cap > 2 * cap0) continue;
            for (int pm = 1; pm >= 0; pm--)
                for (int fl = 1; fl >= 0; fl--) {
                    LaunchShape L = make_shape(c, p, sh[0], sh[1], fl, pm, b->tune_cache != 0, b->grid != nullptr);
                    if (L.block) cand.push_back(L);
                }
        }
        float best = 1e30f;
        LaunchShape win;
        win.block = 0;
        for (size_t q = 0; q < cand.size() && rc2 == SMR_OK; q++) {
            ClassPlan tp = p;
            tp.shape = cand[q];
            float ms = 0.f, ms2 = 0.f;
            rc2 = trial(tp, ms);
            if (rc2 == SMR_OK) rc2 = trial(tp, ms2);   // the first run also warms caches
            if (rc2 == SMR_OK && std::min(ms, ms2) < best) {
                best = std::min(ms, ms2);
                win = cand[q];
            }
        }
        if (rc2 != SMR_OK) break;
        if (win.block) {
            TunedShape &t = c->tuned[p.cls];
            t.valid = true;
            t.block = win.block;
            t.apt = win.apt;
            t.fint_lds = win.fint_lds;
            t.pair_mode = win.pair_mode;
        }
    }
    // restore the batch whatever happened above
    (void)hipMemcpyAsync(b->pos.ptr, kpos.ptr, apos_b, hipMemcpyDeviceToDevice, st);
    (void)hipMemcpyAsync(b->vel.ptr, kvel.ptr, apos_b, hipMemcpyDeviceToDevice, st);
    (void)hipMemcpyAsync(b->state.ptr, kst.ptr, st_b, hipMemcpyDeviceToDevice, st);
    // the trials ran without cell keys and left the stencils of their own
    // last cells in the stored-stencil buffer: every stored stencil is void
    if (b->stencil_key.ptr)
        (void)hipMemsetAsync(b->stencil_key.ptr, 0xff, sizeof(int) * (size_t)b->total_atoms, st);
    hipError_t e = hipStreamSynchronize(st);
    if (rc2 != SMR_OK) return rc2;
    if (e != hipSuccess) return hip_fail(e, "smr_batch_autotune: restore");
    refresh_shapes(b);
    return SMR_OK;
}

extern "C" int smr_batch_set_order(smr_batch *b, int energy_order) {
    if (!b) return fail(SMR_ERR_INVALID, "smr_batch_set_order: NULL argument");
    if (energy_order != 0 && energy_order != 1) return fail(SMR_ERR_INVALID, "smr_batch_set_order: mode must be 0 or 1");
    b->energy_order = energy_order;
    if (!energy_order) {
        HIPCHK(hipSetDevice(b->ctx->device));
        HIPCHK(hipMemcpy(b->work.ptr, b->work_host.data(), sizeof(int) * b->work_host.size(), hipMemcpyHostToDevice));
    }
    return SMR_OK;
}

extern "C" int smr_batch_set_rerank(smr_batch *b, int steps) {
    if (!b) return fail(SMR_ERR_INVALID, "smr_batch_set_rerank: NULL argument");
    if (steps < -1) return fail(SMR_ERR_INVALID, "smr_batch_set_rerank: steps must be >= -1");
    b->rerank_steps = steps;
    return SMR_OK;
}

extern "C" int smr_batch_set_layer_persistent(smr_batch *b, int on, double skin_A) {
    if (!b) return fail(SMR_ERR_INVALID, "smr_batch_set_layer_persistent: NULL argument");
    if (!(skin_A <= 0.0) && !(skin_A < 10.0))
        return fail(SMR_ERR_INVALID, "smr_batch_set_layer_persistent: skin_A = " + std::to_string(skin_A) +
                                         " (need < 10 A; <= 0 keeps the current skin)");
    b->layer_persistent = on != 0;
    if (skin_A > 0.0) b->layer_skin = (float)skin_A;
    return SMR_OK;
}

extern "C" int smr_batch_layer_path(const smr_batch *b, int *persistent, int *fallback, int *rebuilds) {
    if (!b) return fail(SMR_ERR_INVALID, "smr_batch_layer_path: NULL argument");
    if (persistent) *persistent = b->last_layer_persistent;
    if (fallback) *fallback = b->last_layer_fallback;
    if (rebuilds) *rebuilds = b->last_layer_rebuilds;
    return SMR_OK;
}

extern "C" int smr_batch_set_graph(smr_batch *b, int use_graph) {
    if (!b) return fail(SMR_ERR_INVALID, "smr_batch_set_graph: NULL argument");
    b->ls.drop_graph();
    b->ls.use_graph = use_graph != 0;
    return SMR_OK;
}

extern "C" int smr_batch_reset(smr_batch *b) {
    if (!b) return fail(SMR_ERR_INVALID, "smr_batch_reset: NULL argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipMemcpyAsync(b->pos.ptr, b->pos0.ptr, sizeof(float4) * b->total_atoms, hipMemcpyDeviceToDevice,
                          b->ctx->stream));
    HIPCHK(hipMemsetAsync(b->vel.ptr, 0, sizeof(float4) * b->total_atoms, b->ctx->stream));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    return SMR_OK;
}

extern "C" int smr_batch_set_launch(smr_batch *b, int block, int stencil_cache, int pair_sweep) {
    if (!b) return fail(SMR_ERR_INVALID, "smr_batch_set_launch: NULL argument");
    if (block != 0 && block != 64 && block != 128 && block != 256 && block != 512)
        return fail(SMR_ERR_INVALID, "smr_batch_set_launch: block " + std::to_string(block) + " not in {0,64,128,256,512}");
    if (stencil_cache < -1 || stencil_cache > 1)
        return fail(SMR_ERR_INVALID, "smr_batch_set_launch: stencil_cache must be -1, 0 or 1");
    if (pair_sweep < -1 || pair_sweep > 1)
        return fail(SMR_ERR_INVALID, "smr_batch_set_launch: pair_sweep must be -1, 0 or 1");
    b->tune_block = block;
    b->tune_cache = stencil_cache;
    b->tune_pairs = pair_sweep;
    refresh_shapes(b);
    return SMR_OK;
}
