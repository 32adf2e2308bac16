// layer_persist.h -- device side of the coupled layer in one cooperative
// launch (KMODE_LAYER_P of k_relax): coherent exchange, the grid barrier, the
// shared FIRE decision and the neighbour staging / Verlet lists.
//
// Every molecule keeps one 64-thread workgroup for the whole launch.  A step:
//   forces of the own atoms (bonded, intramolecular pairs, surface, and the
//   intermolecular pairs of the atom's list over the staged neighbours)
//   -> v += F dt, the molecule's FIRE record
//   -> publish: positions, velocities, forces and the record, sc1 stores
//   -> grid barrier
//   -> every workgroup reduces all records in the same fixed order and takes
//      the same FIRE decision (relax_pentacene_uff.rs:162-181), then moves
//      its atoms and the staged neighbour atoms with it (the move of a
//      neighbour is recomputed from its published p, v, F: no second barrier)
//   -> lists are rebuilt by every workgroup at the same step, when the
//      largest displacement since the last build, plus a bound on this
//      step's move, exceeds half the skin.
// The exchange buffers are double-buffered by step parity: a workgroup
// writes parity k+1 only after the barrier of step k, which every workgroup
// reaches after its last read of parity k-1.
//
// Coherence across XCDs: the exchange goes through buffer loads and stores
// with the sc1 policy (device-coherent), and the barrier waits for this
// wave's stores before it arrives; no L2 write-back or invalidate is needed
// (those cost ~16 us per barrier at 512 workgroups, the hierarchical
// barrier itself ~3 us).
#pragma once
#include "smr_device.h"
#include "smr_kernels.h"

namespace smr {
namespace lp {

constexpr int kSC1 = 16;                 // cache-policy bits of a buffer op: sc1
constexpr unsigned kSpinLimit = 1u << 24;  // polls (s_sleep 1 each) before a barrier gives up
constexpr int kGroups = kLpGroups;

typedef __attribute__((ext_vector_type(4))) unsigned u32x4;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float4 ldc4(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kSC1));
}
__device__ __forceinline__ void stc4(__amdgpu_buffer_rsrc_t r, unsigned off, float4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, off, 0, kSC1);
}
__device__ __forceinline__ float4 d2f4(double a, double b) { return __builtin_bit_cast(float4, make_double2(a, b)); }
__device__ __forceinline__ double2 f42d(float4 v) { return __builtin_bit_cast(double2, v); }

__device__ __forceinline__ unsigned ld_relaxed(unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// error word: the first error code wins
__device__ __forceinline__ void set_error(const LayerPersist &P, int code) {
    unsigned *w = P.bar + (kGroups + 1) * 32;
    unsigned expected = 0u;
    __hip_atomic_compare_exchange_strong(w, &expected, (unsigned)code, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool has_error(const LayerPersist &P) { return ld_relaxed(P.bar + (kGroups + 1) * 32) != 0u; }

// Grid barrier of generation gen (1, 2, ...) among the launch's workgroups
// (one wave each, all resident: cooperative launch).  Arrivals are counted
// in kGroups counters (workgroup b in group b % kGroups); the last arrival of
// a group bumps the top counter, on which every wave waits.  The wait is
// bounded: it gives up (false, error LP_ERR_BARRIER) after kSpinLimit polls
// or when another wave has recorded an error, so every wave leaves.
__device__ __forceinline__ bool barrier(const LayerPersist &P, unsigned gen, int lane) {
    __builtin_amdgcn_s_waitcnt(0);       // this wave's coherent stores have completed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    int ok = 1;
    if (lane == 0) {
        const unsigned G = gridDim.x, b = blockIdx.x, x = b % kGroups;
        const unsigned nx = (G - x + kGroups - 1) / kGroups;
        const unsigned ngroups = G < (unsigned)kGroups ? G : (unsigned)kGroups;
        unsigned *top = P.bar + kGroups * 32;
        const unsigned old = __hip_atomic_fetch_add(P.bar + x * 32, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1u == gen * nx) __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (ld_relaxed(top) < gen * ngroups) {
            if (++spins > kSpinLimit) {
                set_error(P, LP_ERR_BARRIER);
                ok = 0;
                break;
            }
            if ((spins & 63u) == 0u && has_error(P)) {
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    ok = __shfl(ok, 0, 64);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return ok != 0;
}

// exclusive prefix sum over the wave
__device__ __forceinline__ int wave_excl_scan(int v, int lane) {
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x - v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// FIRE state of the layer, the same in every workgroup
struct Fire {
    double dt, alpha, fmax;
    int npos, steps, status;
    float a, b, dt_move;
    int reset;
};

// The shared FIRE decision from the layer's sums (k_layer_decide's formulas).
__device__ __forceinline__ void decide(Fire &S, const FireConst &P, double vf, double vv, double ff, double fm) {
    const double dt_old = S.dt;
    const bool nonfinite = !(isfinite(vf) && isfinite(vv) && isfinite(ff));
    if (vf > 0.0 && ff > 1e-30 && vv > 1e-30) {
        S.a = (float)(1.0 - S.alpha);
        S.b = (float)(S.alpha * sqrt(vv) / sqrt(ff));
        S.reset = 0;
        S.npos++;
        if (S.npos > P.n_min) {
            S.dt = fmin(S.dt * P.f_inc, P.dt_max);
            S.alpha *= P.f_alpha;
        }
    } else {
        S.npos = 0;
        S.dt *= P.f_dec;
        S.alpha = P.alpha0;
        S.a = S.b = 0.f;
        S.reset = 1;
    }
    S.dt_move = (float)dt_old;
    S.steps++;
    S.fmax = sqrt(fm);
    if (nonfinite) S.status = SMR_SYS_NONFINITE;
    else if (S.fmax < P.fconv) S.status = SMR_SYS_CONVERGED;
    else if (S.steps >= P.max_steps) S.status = SMR_SYS_MAXSTEPS;
}

// The move of one atom under the decision: v = a v + b F (0 on a reset),
// p += v dt.  Every workgroup that holds the atom (its own, and those that
// stage it as a neighbour) computes it with these exact operations.
__device__ __forceinline__ float3 mixed_velocity(float3 v, float3 f, float a, float b, int reset) {
    if (reset) return make_float3(0.f, 0.f, 0.f);
    return make_float3(__builtin_fmaf(v.x, a, f.x * b), __builtin_fmaf(v.y, a, f.y * b), __builtin_fmaf(v.z, a, f.z * b));
}
__device__ __forceinline__ float3 moved(float3 p, float3 v, float dt) {
    return make_float3(__builtin_fmaf(v.x, dt, p.x), __builtin_fmaf(v.y, dt, p.y), __builtin_fmaf(v.z, dt, p.z));
}

// LDS staging area (at byte offset P.lds_off of the dynamic LDS)
struct Stage {
    float4 *pos;    // [maxs] moved neighbour positions (local frame of their molecule), w = instance (int bits)
    float4 *req;    // [maxs]
    int *gid;       // [maxs] global atom index
    float4 *off;    // [kLpMaxInst] frame offset of the instance (xyz), w = first staged index (int bits)
    float4 *sph;    // [kLpMaxInst] instance centre in the own frame, radius
    int *nat;       // [kLpMaxInst] atoms of the instance
    int *mol;       // [kLpMaxInst] the instance's molecule (system index)
    int *misc;      // [4] instances, staged atoms, list chunks
    int *choff;     // [64] first chunk of each own atom
    int *lcnt;      // [64] list length of each own atom
    float4 *ipart;  // [lds_atoms * chunks] partial force of each chunk
    unsigned char *iatom;   // [lds_atoms * chunks] own atom of each chunk
    unsigned short *nl;     // [maxl][lds_atoms] list entries: staged atom (bits 0-9), instance (bits 10-15)
};
__device__ __forceinline__ Stage stage_of(unsigned char *smem, const LayerPersist &P, int lds_atoms) {
    Stage S;
    unsigned char *b = smem + P.lds_off;
    S.pos = reinterpret_cast<float4 *>(b);
    S.req = S.pos + P.maxs;
    S.off = S.req + P.maxs;
    S.sph = S.off + kLpMaxInst;
    S.gid = reinterpret_cast<int *>(S.sph + kLpMaxInst);
    S.nat = S.gid + P.maxs;
    S.mol = S.nat + kLpMaxInst;
    S.misc = S.mol + kLpMaxInst;
    S.choff = S.misc + 4;
    S.lcnt = S.choff + 64;
    S.ipart = reinterpret_cast<float4 *>(S.lcnt + 64);
    const int items = lds_atoms * lp_chunks(P.maxl);
    S.iatom = reinterpret_cast<unsigned char *>(S.ipart + items);
    S.nl = reinterpret_cast<unsigned short *>(S.iatom + ((items + 15) & ~15));
    return S;
}

}  // namespace lp
}  // namespace smr
