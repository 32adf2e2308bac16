// plain_noalloc.hip -- the plain method with no device allocation: the
// scratch of its kernels (hneigh, force pieces, per-atom energies and the
// reduction output) lives in a fixed file-scope device array of each GPU
// instead of buffers the library allocates on the batch.
//
// The array holds one system of up to SMR_MAX_ATOMS (1024) atoms and
// kMaxPieces force pieces (4 per dihedral and inversion, 3 per angle): 600 KiB
// per GPU, held until the process ends.  Its device address is resolved once
// per device by smr_ctx_create (noalloc_setup), with that device current.
// Every call on one GPU uses the same array, so calls that overlap on one GPU
// race; nothing in it needs to start at zero (each step writes every entry it
// reads).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/surfmol_relax.h"
#include "smr_internal.h"

namespace {
constexpr int kMaxAtoms = SMR_MAX_ATOMS;
constexpr int kMaxPieces = 32768;
constexpr size_t kHneighBytes = sizeof(float4) * 4 * kMaxAtoms;
constexpr size_t kFintBytes = sizeof(float4) * kMaxPieces;
constexpr size_t kEatomBytes = sizeof(float) * kMaxAtoms;
constexpr size_t kEsurfBytes = sizeof(float) * kMaxAtoms;
constexpr size_t kEpairBytes = sizeof(double) * kMaxAtoms;
constexpr size_t kCvfBytes = sizeof(double) * 16;
constexpr size_t kScratchBytes = kHneighBytes + kFintBytes + kEatomBytes + kEsurfBytes + kEpairBytes + kCvfBytes;
static_assert(kScratchBytes <= 600 * 1024, "fixed scratch larger than stated");
constexpr int kMaxDevices = 64;

std::mutex g_mu;
char *g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch` per GPU
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace smr {
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return SMR_OK;   // the entry point reports it
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch_ptr[device]) return SMR_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "smr_ctx_create: address of the fixed plain-method scratch");
    g_scratch_ptr[device] = static_cast<char *>(p);
    return SMR_OK;
}
}  // namespace smr

extern "C" int smr_relax_plain_noalloc(smr_batch *b, int isys, const smr_fire_params *fire, int max_steps, double fconv,
                                       smr_relax_result *result) {
    using namespace smr;
    if (!b || !result) return fail(SMR_ERR_INVALID, "smr_relax_plain_noalloc: NULL argument");
    const int dev = b->ctx->device;
    char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (dev >= 0 && dev < kMaxDevices) base = g_scratch_ptr[dev];
    }
    if (!base)
        return fail(SMR_ERR_STATE, "smr_relax_plain_noalloc: no fixed scratch on device " + std::to_string(dev));
    PlainScratch S;
    S.hneigh = reinterpret_cast<float4 *>(base);
    S.fint = reinterpret_cast<float4 *>(base + kHneighBytes);
    S.eatom = reinterpret_cast<float *>(base + kHneighBytes + kFintBytes);
    S.esurf = reinterpret_cast<float *>(base + kHneighBytes + kFintBytes + kEatomBytes);
    S.epair = reinterpret_cast<double *>(base + kHneighBytes + kFintBytes + kEatomBytes + kEsurfBytes);
    S.cvf = reinterpret_cast<double *>(base + kHneighBytes + kFintBytes + kEatomBytes + kEsurfBytes + kEpairBytes);
    S.max_atoms = kMaxAtoms;
    S.max_pieces = kMaxPieces;
    return relax_plain_run(b, isys, fire, max_steps, fconv, result, &S, "smr_relax_plain_noalloc");
}
