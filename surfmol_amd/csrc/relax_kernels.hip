// relax_kernels.hip -- HIP kernels (gfx950) of the multi-system relaxation path.
//
//  k_relax      the fused persistent kernel: one workgroup owns one system
//               at a time and runs bonds -> angles/dihedrals/inversions ->
//               per-atom gather + pairs + surface -> FIRE (or damped MD) for
//               many steps without leaving the kernel.  Positions, REQ,
//               hneigh and the bonded force pieces live in LDS; per-system
//               reductions are wave shuffles (+ LDS across waves) in FP64.
//               Forces only: integration needs no energy, and the energies
//               reported at the end are evaluated in FP64 by k_energy64.
//               Systems are taken from a work list through an atomic
//               counter, so uneven systems balance across the chip.
//  k_plain_*    the plain method: one system, one kernel per force term per
//               step, FIRE decided on the host (the reference's shape,
//               opencl/UFF.cl:141-1011 driven by a host loop).
//
// Per-term math is in smr_device.h and shared by both paths.
// this translation unit holds the pair sweep's schedule table (kSched)
#define SMR_SWEEP_SCHEDULE_TABLE
#include "smr_device.h"
#include "smr_kernels.h"
#include "layer_persist.h"

namespace smr {

// Workgroup barrier of the fused kernel.  A one-wave workgroup needs no
// hardware barrier: its lanes run in lockstep and its LDS accesses complete
// in issue order, so a wavefront-scope fence (a code-motion barrier that
// emits no waits) orders the phases; wider workgroups use s_barrier.
template <int NW>
__device__ __forceinline__ void block_sync() {
    if constexpr (NW == 1) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

// The launch's arguments in the kernel-argument segment (constant address
// space: scalar loads).  fresh_args() hides the pointer from the optimiser,
// so that fields read through it are loaded where they are used instead of
// being held in scalar registers for the whole kernel (which spills them).
typedef const __attribute__((address_space(4))) KernelArgs ConstArgs;
__device__ __forceinline__ ConstArgs *fresh_args() {
    ConstArgs *p = (ConstArgs *)(__builtin_amdgcn_kernarg_segment_ptr());
    asm volatile("" : "+s"(p));
    return p;
}

// Value of lane (lane+1) mod 64 (DPP wave_rol:1).
__device__ __forceinline__ float wave_rol1(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x134, 0xf, 0xf, true));
}

// Value of lane (lane+1) mod (last+1) for lanes 0..last: the first last+1
// lanes rotated by one towards lane 0 (DPP wave_shl:1, then lane `last`
// takes lane 0's value).  Lanes above `last` receive garbage.
__device__ __forceinline__ float wave_rot1(float x, int last, int lane) {
    const int xi = __float_as_int(x);
    const int y = __builtin_amdgcn_update_dpp(0, xi, 0x130, 0xf, 0xf, true);
    const int x0 = __builtin_amdgcn_readlane(xi, 0);
    return __int_as_float(lane == last ? x0 : y);
}

// Refresh the register stencils of the atoms of one wave that entered a new
// cell (need = true; the wave calls this converged).  The changed atoms are
// served nslot (<= 4) at a time: each of 16 lanes loads one 4-node z column of
// an atom's 4x4x4 stencil, contracts it with that atom's PLQH and passes it
// through the wave's LDS staging area (16*nslot float4), from which the
// atom's own lane picks up its 16 values.  A wave then waits for one round of
// grid loads per nslot changed atoms instead of one per x plane of every
// changed atom; the values equal grid_contract_stencil's.
__device__ __forceinline__ void refresh_stencils(const float4 *grid, int gnx, int gny, int gnz, bool need,
                                                 const GridCell &gc, int key,
                                                 float4 plq, float4 (&rs)[16], int &rkey, float4 *stage, int nslot,
                                                 int lane) {
    unsigned long long pend = __builtin_amdgcn_ballot_w64(need);
    const int q = lane >> 4, col = lane & 15;
    while (pend) {
        // owner of slot q: the q-th pending lane, if any
        const unsigned long long m1 = pend & (pend - 1), m2 = m1 & (m1 - 1), m3 = m2 & (m2 - 1);
        const unsigned long long mq = q == 0 ? pend : q == 1 ? m1 : q == 2 ? m2 : m3;
        const bool serve = q < nslot && mq != 0ull;
        const int owner = serve ? __builtin_ctzll(mq) : lane;
        const int oix = __shfl(gc.ix, owner, 64), oiy = __shfl(gc.iy, owner, 64), oiz = __shfl(gc.iz, owner, 64);
        const float4 op = make_float4(__shfl(plq.x, owner, 64), __shfl(plq.y, owner, 64), __shfl(plq.z, owner, 64), 0.f);
        if (serve) {
            int xi = oix + (col >> 2); if (xi >= gnx) xi -= gnx;
            int yi = oiy + (col & 3); if (yi >= gny) yi -= gny;
            const float4 *c = grid + ((size_t)(xi * gny + yi) * gnz + oiz);
            stage[q * 16 + col] = make_float4(dotplq(op, c[0]), dotplq(op, c[1]), dotplq(op, c[2]), dotplq(op, c[3]));
        }
        block_sync<1>();
        const int nserved = min(nslot, __builtin_popcountll(pend));
        const int rank = __builtin_popcountll(pend & ((1ull << lane) - 1ull));
        if (need && rank < nserved) {
#pragma unroll
            for (int e = 0; e < 16; e++) rs[e] = stage[rank * 16 + e];
            rkey = key;
            need = false;
        }
        block_sync<1>();
        for (int t = 0; t < nserved; t++) pend &= pend - 1ull;
    }
}

// ===========================================================================
// fused persistent kernel
// ===========================================================================
// One atom per thread: at most 256 registers, so that two waves share each
// SIMD (the LDS of a <= 64-atom system admits two per SIMD as well); wider
// shapes keep the compiler's choice (their per-thread atom arrays would
// spill).  LAYER: the launch carries a coupled layer's cell list (its
// intermolecular partial forces and, in KMODE_LAYER, the shared FIRE
// state); the code for it is compiled into these instances only.
// PERSIST: a coupled layer in one cooperative launch (KMODE_LAYER_P, one
// workgroup per molecule for all steps; layer_persist.h).  Its workgroups
// are one per molecule and at most a few per CU, so it is built for one wave
// per SIMD (all registers).
template <int BLOCK, int APT, bool FINT_LDS, bool LAYER, bool PERSIST = false>
__global__ __launch_bounds__(BLOCK, (APT == 1 && !PERSIST) ? 2 : 1) void k_relax(KernelArgs A) {
    constexpr int NW = BLOCK / 64;
    // one atom per thread: each atom's surface stencil lives in registers
    constexpr bool RSTEN = (APT == 1);
    // one wave, one atom per thread: the pair sweep carries reaction forces
    // in registers that move one lane per sweep
    constexpr bool ROT = (BLOCK == 64 && APT == 1);
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const DevTables &T = A.tab;

    // LDS carve-up (sizes fixed per launch by the host, relax_lds_bytes)
    const int rows = relax_lds_rows(BLOCK, APT, A.lds_atoms);
    float4 *s_pos = reinterpret_cast<float4 *>(smem);        // [rows] (ROT: rows 0..N/2 again after row N-1)
    float4 *s_req = s_pos + rows;                            // [rows]
    float4 *s_h = s_req + rows;                              // [4*natoms] hneigh
    float4 *s_fp = s_h + 4 * A.lds_atoms;                    // [NWP][natoms] pair reaction forces (not ROT)
    const int NWP = ROT ? 0 : (A.pair_mode ? NW : 1);        // one reaction buffer per wave
    double *s_red = reinterpret_cast<double *>(s_fp + NWP * A.lds_atoms);  // [NW*8]
    int *s_next = reinterpret_cast<int *>(s_red + NW * 8);
    // interaction slots and piece destinations: staged in LDS, except for
    // the widest shapes (4 atoms per thread, up to 1024 atoms), whose tables
    // are read through the caches instead so that the atom data fits
    constexpr bool TABLES_LDS = (APT < 4);
    const int tn_d = TABLES_LDS ? A.lds_nd : 0, tn_i = TABLES_LDS ? A.lds_ni : 0, tn_a = TABLES_LDS ? A.lds_na : 0;
    uint2 *s_dih = reinterpret_cast<uint2 *>(s_next + 4);                // packed slots
    uint2 *s_inv = s_dih + tn_d;
    uint2 *s_dihd = s_inv + tn_i;                                       // piece destinations
    uint2 *s_invd = s_dihd + tn_d;                                      // (not ROT)
    uint2 *s_angd = s_invd + (ROT ? 0 : tn_i);
    unsigned *s_ang = reinterpret_cast<unsigned *>(s_angd + tn_a);
    float *s_fx = reinterpret_cast<float *>(s_ang + tn_a);
    const uint2 *t_dih, *t_inv, *t_dihd, *t_invd, *t_angd;
    const unsigned *t_ang;
    float *fint_x, *fint_y, *fint_z;

    // a coupled layer that has converged (or stopped) does no more work
    if (LAYER && (A.mode == KMODE_LAYER || A.mode == KMODE_LAYER_P) && A.layer->status != SMR_SYS_RUNNING) return;
    // a launch of one step (coupled-layer steps, force evaluations) with a
    // workgroup for every system: each takes its own, without a contended
    // counter; otherwise systems are fetched through the atomic counter in
    // the order the workgroups start (concurrent launches share the chip,
    // so not every workgroup starts at once)
    const bool one_each = PERSIST || (A.steps_per_launch == 1 && A.nwork <= (int)gridDim.x);
    int pass = 0;
    while (true) {
        // ---- fetch the next system from the work list ----
        // (tables and state arrays read through fresh_args(): these reads
        // happen once per system, and the step loop keeps its scalar
        // registers for its own constants)
        ConstArgs *fa = fresh_args();
        int wi;
        if (one_each) {
            if (pass++) break;
            wi = blockIdx.x;
        } else {
            if (tid == 0) s_next[0] = atomicAdd(fa->work_counter, 1);
            block_sync<NW>();
            wi = s_next[0];
            block_sync<NW>();
        }
        if (wi >= fa->nwork) break;
        const int s = fa->work_list[wi];
        const SysDesc sd = fa->tab.sys[s];
        const TopoDesc td = fa->tab.topo[sd.topo];
        SysState st = fa->state[s];
        // a finished system is not stepped again (forces are always evaluated)
        // (a persistent layer's workgroups all take part in every step)
        if (!PERSIST && A.mode != KMODE_EVAL && st.status != SMR_SYS_RUNNING) continue;
        const int N = td.natoms;
        const int abase = sd.atom_base;
        if (FINT_LDS) {
            fint_x = s_fx;
        } else {
            fint_x = fa->fint_scratch + (size_t)blockIdx.x * 3 * fa->nf_max;
        }
        fint_y = fint_x + fa->nf_max;
        fint_z = fint_y + fa->nf_max;
        // ROT with pieces in LDS: the gather masks by weight (0 or 1), so
        // every float it may read must be finite: the slots this system does
        // not write ([nf, nf_max) of each array) and the tail are zeroed
        if constexpr (ROT && FINT_LDS) {
            for (int q = td.nf + tid; q < fa->nf_max; q += BLOCK) fint_x[q] = fint_y[q] = fint_z[q] = 0.f;
            if (tid < kPieceTail) fint_z[fa->nf_max + tid] = 0.f;
        }

        // ---- stage the system ----
        int4 my_ng[APT];
        float4 my_bk[APT], my_bl[APT], my_plq[APT];
        int2 my_a2f[APT];
        float3 my_v[APT];
#pragma unroll
        for (int k = 0; k < APT; k++) {
            const int i = tid + k * BLOCK;
            if (i < N) {
                const int ta = td.atom_off + i;
                s_pos[i] = fa->pos[abase + i];
                s_req[i] = fa->tab.req[ta];
                my_ng[k] = fa->tab.ng[ta];
                my_bk[k] = fa->tab.bk[ta];
                my_bl[k] = fa->tab.bl[ta];
                my_plq[k] = fa->tab.plq[ta];
                my_a2f[k] = fa->tab.a2f[ta];
                float4 v = fa->vel[abase + i];
                my_v[k] = make_float3(v.x, v.y, v.z);
            } else {
                my_ng[k] = make_int4(-1, -1, -1, -1);
                my_bk[k] = my_bl[k] = my_plq[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                my_a2f[k] = make_int2(0, 0);
                my_v[k] = make_float3(0.f, 0.f, 0.f);
            }
        }
        // ROT: the pair sweep's lane schedule (sweep_schedules.h) for up to
        // kSchedMaxAtoms atoms: the lane takes atom s_av (-1: none) and meets
        // rows lane+1 .. lane+s_T; bit t-1 of s_skip drops round t's pair (not
        // the meeting that sums it, or a 1-2 neighbour)
        // s_src: the atom's further pair-force sums (one byte each), s_cp: the
        // atoms copied into rows N + lane and N + lane + 64 (bits 0-7, 8-15;
        // bit 16: the second row exists)
        [[maybe_unused]] int s_av = -1, s_T = 0;
        [[maybe_unused]] unsigned s_skip = ~0u, s_src = 0xffffffffu, s_cp = 0u;
        if constexpr (ROT) {
            if (N <= kSchedMaxAtoms) {
                const SweepSchedule &SS = kSched[N];
                s_T = SS.T;
                s_av = SS.a[lane];
                s_skip = ~SS.keep[lane];
                if (tid < N) s_src = SS.src[lane];
                const int x1 = N + lane, x2 = x1 + 64;
                const int j1 = SS.g[x1 & 63], j2 = SS.g[x2 & 63];
                s_cp = (unsigned)(j1 >= 0 ? j1 : 0) | ((unsigned)(j2 >= 0 ? j2 : 0) << 8) | (x2 < 64 + s_T ? 1u << 16 : 0u);
                if (s_av >= 0) {
                    const int4 ng = s_av == tid ? my_ng[0] : fa->tab.ng[td.atom_off + s_av];
                    for (int t = 1; t <= s_T; t++) {
                        const int j = SS.g[(lane + t) & 63];
                        if (j >= 0 && (j == ng.x || j == ng.y || j == ng.z || j == ng.w)) s_skip |= 1u << (t - 1);
                    }
                }
            }
        }
        // ROT: copies of atoms in the rows after row N-1, so that the partner
        // of a lane's round t is row lane + t: the schedule's rows (an empty
        // slot holds atom 0, whose meetings there are all dropped), or rows
        // 0..N/2 again for more atoms.  Called after the atom rows changed
        // (staging: with the REQ rows; every move: positions).
        auto copy_rows = [&](bool with_req) {
            if constexpr (ROT) {
                if (N <= kSchedMaxAtoms) {
                    block_sync<1>();
                    const int x1 = N + lane, j1 = (int)(s_cp & 0xffu), j2 = (int)((s_cp >> 8) & 0xffu);
                    if (x1 < 64 + s_T) {
                        s_pos[x1] = s_pos[j1];
                        if (with_req) s_req[x1] = s_req[j1];
                    }
                    if (s_cp >> 16) {
                        s_pos[x1 + 64] = s_pos[j2];
                        if (with_req) s_req[x1 + 64] = s_req[j2];
                    }
                } else if (tid <= (N >> 1)) {
                    s_pos[N + tid] = s_pos[tid];
                    if (with_req) s_req[N + tid] = s_req[tid];
                }
            }
        };
        copy_rows(true);
        // ROT: the gather of each step follows a plan fixed per system.  An
        // atom's force pieces are summed in rounds of eight; the wave runs
        // as many rounds as its heaviest atom needs.  Lanes without an atom
        // take the tail of the heaviest atoms' ranges (ranked by piece count,
        // then index): the atom sums its first 8 ceil(r/2) pieces (r = its
        // rounds), the helper lane the rest, and the atom adds the helper's
        // sum after its own.  (g_start, g_cnt): the lane's range; g_help: the
        // helper lane of the lane's atom, or -1.
        [[maybe_unused]] int g_start = 0, g_cnt = 0, g_help = -1;
        if constexpr (ROT) {
            const bool own = tid < N;
            const int c = own ? my_a2f[0].y : 0;
            const int r = (c + 7) >> 3;
            int rank = 0;
            for (int j = 0; j < N; j++) {
                const int cj = __shfl(c, j, 64);
                rank += (cj > c || (cj == c && j < tid)) ? 1 : 0;
            }
            const int nhelp = 64 - N;
            const bool helped = own && r >= 2 && rank < nhelp;
            const int a = helped ? 8 * ((r + 1) >> 1) : c;
            g_help = helped ? N + rank : -1;
            int hatom = -1;
            for (int k = 0; k < nhelp; k++) {
                const unsigned long long m = __builtin_amdgcn_ballot_w64(helped && rank == k);
                if (m != 0ull && lane == N + k) hatom = __builtin_ctzll(m);
            }
            const int src = hatom >= 0 ? hatom : lane;
            const int s_x = __shfl(my_a2f[0].x, src, 64), s_a = __shfl(a, src, 64), s_c = __shfl(c, src, 64);
            g_start = own ? my_a2f[0].x : (hatom >= 0 ? s_x + s_a : 0);
            g_cnt = own ? a : (hatom >= 0 ? s_c - s_a : 0);
        }
        // interaction slots and piece destinations stay in LDS for all steps
        if constexpr (TABLES_LDS) {
            for (int q = tid; q < td.nangles; q += BLOCK) {
                s_ang[q] = fa->tab.ang_ng[td.ang_off + q];
                s_angd[q] = fa->tab.ang_dst[td.ang_off + q];
            }
            for (int q = tid; q < td.ndih; q += BLOCK) {
                s_dih[q] = fa->tab.dih_ng[td.dih_off + q];
                s_dihd[q] = fa->tab.dih_dst[td.dih_off + q];
            }
            for (int q = tid; q < td.ninv; q += BLOCK) {
                s_inv[q] = fa->tab.inv_ng[td.inv_off + q];
                if constexpr (!ROT) s_invd[q] = fa->tab.inv_dst[td.inv_off + q];
            }
            t_ang = s_ang; t_dih = s_dih; t_inv = s_inv; t_angd = s_angd; t_dihd = s_dihd;
            // one wave, one atom per thread: read through the caches
            t_invd = ROT ? fa->tab.inv_dst + td.inv_off : s_invd;
        } else {
            t_ang = fa->tab.ang_ng + td.ang_off; t_dih = fa->tab.dih_ng + td.dih_off; t_inv = fa->tab.inv_ng + td.inv_off;
            t_angd = fa->tab.ang_dst + td.ang_off; t_dihd = fa->tab.dih_dst + td.dih_off; t_invd = fa->tab.inv_dst + td.inv_off;
        }
        block_sync<NW>();

        // cell of each own atom's stored surface stencil (fa->stencil): the
        // PLQH-contracted 4x4x4 stencil is re-read from the grid only when
        // the atom enters another cell
        int skey[APT];
#pragma unroll
        for (int k = 0; k < APT; k++) {
            const int i = tid + k * BLOCK;
            skey[k] = (fa->stencil_key && i < N) ? fa->stencil_key[abase + i] : -1;
        }
        // one atom per thread: the contracted stencil is held in registers
        // for the whole launch (rkey = its cell, -1: none); the stored copy
        // is read once here and written back when the system leaves
        float4 rs[RSTEN ? 16 : 1];
        int rkey = -1;
        if constexpr (RSTEN) {
            if (fa->stencil && tid < N && skey[0] != -1) {
                rkey = skey[0];
                if (rkey >= 0) {
                    const float4 *sc = fa->stencil + (size_t)abase * 16 + tid;
#pragma unroll
                    for (int e = 0; e < 16; e++) rs[e] = sc[e * N];
                }
            }
        }
        auto persist_stencil = [&]() {
            if constexpr (RSTEN) {
                if (fa->stencil && tid < N && rkey != skey[0]) {
                    if (rkey >= 0) {
                        float4 *sc = fa->stencil + (size_t)abase * 16 + tid;
#pragma unroll
                        for (int e = 0; e < 16; e++) sc[e * N] = rs[e];
                    }
                    skey[0] = rkey;
                }
            }
        };
        const int nsteps = A.steps_per_launch;
        int step_in_launch = 0;

        // ---- coupled layer in one cooperative launch (KMODE_LAYER_P) ----
        // The shared FIRE state, the staging area of the neighbour atoms and
        // the own atom's Verlet list (length lcnt, built at pref) persist
        // over the launch's steps.
        [[maybe_unused]] lp::Fire fs{};
        [[maybe_unused]] int lcnt = 0, kstep = 0, nrebuild = 0;
        [[maybe_unused]] unsigned gen = 0;
        [[maybe_unused]] bool lp_failed = false;
        [[maybe_unused]] float3 pref = make_float3(0.f, 0.f, 0.f);
        [[maybe_unused]] float4 msph = make_float4(0.f, 0.f, 0.f, 0.f);
        [[maybe_unused]] float lp_rc = 0.f;
        if constexpr (PERSIST) lp_rc = sqrtf(A.cells.rc2);
        const LayerPersist &LP = A.lp;
        const lp::Stage stg = lp::stage_of(smem, LP, A.lds_atoms);
        const unsigned nat_all = (unsigned)LP.natoms;
        const unsigned nmol = gridDim.x;
        const auto rx = lp::rsrc(LP.xbuf, 96u * nat_all);
        const auto rrec = lp::rsrc(LP.xrec, 2u * nmol * kLpRec * 8u);
        // bounding sphere of the own atoms (origin-grid frame)
        auto own_sphere = [&]() -> float4 {
            const bool own = tid < N;
            float3 g = make_float3(0.f, 0.f, 0.f);
            if (own) {
                const float4 p = s_pos[tid];
                g = make_float3(p.x + (float)sd.gx * T.ostep_x, p.y + (float)sd.gy * T.ostep_y, p.z + (float)sd.gz * T.ostep_z);
            }
            float lo[3] = {own ? g.x : INFINITY, own ? g.y : INFINITY, own ? g.z : INFINITY};
            float hi[3] = {own ? g.x : -INFINITY, own ? g.y : -INFINITY, own ? g.z : -INFINITY};
#pragma unroll
            for (int o = 32; o > 0; o >>= 1)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    lo[c] = fminf(lo[c], __shfl_xor(lo[c], o, 64));
                    hi[c] = fmaxf(hi[c], __shfl_xor(hi[c], o, 64));
                }
            const float3 cc = make_float3(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]));
            const float3 d = g - cc;
            float r2 = own ? dot3(d, d) : 0.f;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) r2 = fmaxf(r2, __shfl_xor(r2, o, 64));
            return make_float4(cc.x, cc.y, cc.z, sqrtf(r2));
        };
        // publish this step's own atoms (position before the move, velocity
        // after v += F dt, force) and the molecule's record
        auto publish = [&](int par, float3 v, float3 f, double r0, double r1, double r2, double r3, double r4, double r5,
                           float4 sph) {
            if (tid < N) {
                const unsigned a = (unsigned)(abase + tid);
                lp::stc4(rx, ((par * 3u + 0u) * nat_all + a) * 16u, s_pos[tid]);
                lp::stc4(rx, ((par * 3u + 1u) * nat_all + a) * 16u, make_float4(v.x, v.y, v.z, 0.f));
                lp::stc4(rx, ((par * 3u + 2u) * nat_all + a) * 16u, make_float4(f.x, f.y, f.z, 0.f));
            }
            if (tid < 4) {
                const float4 w = tid == 0 ? lp::d2f4(r0, r1) : tid == 1 ? lp::d2f4(r2, r3) : tid == 2 ? lp::d2f4(r4, r5) : sph;
                lp::stc4(rrec, (((unsigned)par * nmol + blockIdx.x) * kLpRec) * 8u + (unsigned)tid * 16u, w);
            }
        };
        // staged neighbour atoms moved under this step's decision (the same
        // operations as their own workgroup's move)
        auto restage = [&](int par) {
            const int total = stg.misc[1];
            for (int j0 = 0; j0 < total; j0 += 4 * 64) {
                float4 P4[4], V4[4], F4[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int j = j0 + u * 64 + lane;
                    if (j < total) {
                        const unsigned g = (unsigned)stg.gid[j];
                        P4[u] = lp::ldc4(rx, ((par * 3u + 0u) * nat_all + g) * 16u);
                        V4[u] = lp::ldc4(rx, ((par * 3u + 1u) * nat_all + g) * 16u);
                        F4[u] = lp::ldc4(rx, ((par * 3u + 2u) * nat_all + g) * 16u);
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int j = j0 + u * 64 + lane;
                    if (j < total) {
                        const float3 v = lp::mixed_velocity(f3(V4[u]), f3(F4[u]), fs.a, fs.b, fs.reset);
                        const float3 p = lp::moved(f3(P4[u]), v, fs.dt_move);
                        stg.pos[j] = make_float4(p.x, p.y, p.z, stg.pos[j].w);
                    }
                }
            }
            block_sync<NW>();
        };
        // neighbour instances (molecule, box image) within reach of the own
        // sphere, their atoms staged, and the own atom's list of staged atoms
        // within cutoff + skin; delta bounds this step's move of any atom
        auto build = [&](int par, float delta) -> bool {
            const CellGrid &C = A.cells;
            const float reach = lp_rc + LP.skin + 2.f * delta + 0.05f;
            int ninst = 0;
            for (int q0 = 0; q0 < (int)nmol; q0 += 64) {
                const int q = q0 + lane;
                unsigned mask = 0u;
                float4 ct = make_float4(0.f, 0.f, 0.f, 0.f);
                int n0x = 0, n0y = 0;
                if (q < (int)nmol) {
                    ct = lp::ldc4(rrec, (((unsigned)par * nmol + q) * kLpRec) * 8u + 48u);
                    if (C.pbc_x) n0x = (int)rintf((msph.x - ct.x) * C.inv_Lx);
                    if (C.pbc_y) n0y = (int)rintf((msph.y - ct.y) * C.inv_Ly);
                    const float lim = msph.w + ct.w + reach;
#pragma unroll
                    for (int ix = 0; ix < 3; ix++)
#pragma unroll
                        for (int iy = 0; iy < 3; iy++) {
                            if ((!C.pbc_x && ix != 1) || (!C.pbc_y && iy != 1)) continue;
                            const int bx = C.pbc_x ? n0x + ix - 1 : 0, by = C.pbc_y ? n0y + iy - 1 : 0;
                            if (q == (int)blockIdx.x && bx == 0 && by == 0) continue;   // the molecule itself
                            const float dx = ct.x + (float)bx * C.Lx - msph.x, dy = ct.y + (float)by * C.Ly - msph.y,
                                        dz = ct.z - msph.z;
                            if (dx * dx + dy * dy + dz * dz < lim * lim) mask |= 1u << (ix * 3 + iy);
                        }
                }
                const int c = __builtin_popcount(mask);
                const int off = lp::wave_excl_scan(c, lane);
                const int tot = lp::wave_sum_i(c);
                if (ninst + tot > kLpMaxInst) {
                    if (lane == 0) lp::set_error(LP, LP_ERR_INST);
                    return false;
                }
                int k = ninst + off;
                while (mask) {
                    const int bit = __builtin_ctz(mask);
                    mask &= mask - 1u;
                    const int ix = bit / 3, iy = bit % 3;
                    const int bx = C.pbc_x ? n0x + ix - 1 : 0, by = C.pbc_y ? n0y + iy - 1 : 0;
                    const int t = A.work_list[q];
                    const SysDesc dt_ = T.sys[t];
                    const float3 dg = make_float3((float)((dt_.gx - sd.gx) + bx * C.nLx) * T.ostep_x,
                                                  (float)((dt_.gy - sd.gy) + by * C.nLy) * T.ostep_y,
                                                  (float)(dt_.gz - sd.gz) * T.ostep_z);
                    stg.off[k] = make_float4(dg.x, dg.y, dg.z, 0.f);
                    stg.sph[k] = make_float4(ct.x + (float)bx * C.Lx - (float)sd.gx * T.ostep_x,
                                             ct.y + (float)by * C.Ly - (float)sd.gy * T.ostep_y,
                                             ct.z - (float)sd.gz * T.ostep_z, ct.w);
                    stg.nat[k] = T.topo[dt_.topo].natoms;
                    stg.mol[k] = t;
                    k++;
                }
                ninst += tot;
            }
            block_sync<NW>();
            // first staged index of each instance
            const int na_k = lane < ninst ? stg.nat[lane] : 0;
            const int first = lp::wave_excl_scan(na_k, lane);
            const int total = lp::wave_sum_i(na_k);
            if (total > LP.maxs) {
                if (lane == 0) lp::set_error(LP, LP_ERR_STAGE);
                return false;
            }
            if (lane < ninst) stg.off[lane].w = __int_as_float(first);
            if (lane == 0) {
                stg.misc[0] = ninst;
                stg.misc[1] = total;
            }
            block_sync<NW>();
            for (int k = 0; k < ninst; k++) {
                const int t = stg.mol[k], n = stg.nat[k], f0 = __float_as_int(stg.off[k].w);
                const SysDesc dt_ = T.sys[t];
                const int aoff = T.topo[dt_.topo].atom_off;
                for (int a = lane; a < n; a += 64) {
                    stg.gid[f0 + a] = dt_.atom_base + a;
                    stg.req[f0 + a] = T.req[aoff + a];
                    stg.pos[f0 + a] = make_float4(0.f, 0.f, 0.f, __int_as_float(k));
                }
            }
            block_sync<NW>();
            restage(par);
            // the own atom's list
            int cnt = 0;
            if (tid < N) {
                const float3 pi = f3(s_pos[tid]);
                const float rl = lp_rc + LP.skin;
                const float rl2 = rl * rl;
                unsigned short *nl = stg.nl + tid;
                for (int k = 0; k < ninst; k++) {
                    const float4 sp = stg.sph[k];
                    const float3 d = pi - f3(sp);
                    const float lim = sp.w + rl + delta + 0.05f;
                    if (dot3(d, d) >= lim * lim) continue;
                    const float4 o = stg.off[k];
                    const int f0 = __float_as_int(o.w), n = stg.nat[k];
                    for (int a = 0; a < n; a++) {
                        const float3 dp = (f3(stg.pos[f0 + a]) - pi) + f3(o);
                        if (dot3(dp, dp) < rl2) {
                            if (cnt < LP.maxl) nl[cnt * A.lds_atoms] = (unsigned short)((f0 + a) | (k << 10));
                            cnt++;
                        }
                    }
                }
                pref = pi;
            }
            lcnt = cnt;
            const bool over = __builtin_amdgcn_ballot_w64(cnt > LP.maxl) != 0ull;
            if (over) {
                if (lane == 0) lp::set_error(LP, LP_ERR_LIST);
                return false;
            }
            // chunks of kLpChunk entries: the wave's lanes take them in turn
            const int nch = (cnt + kLpChunk - 1) / kLpChunk;
            const int co = lp::wave_excl_scan(nch, lane);
            stg.choff[lane] = co;
            stg.lcnt[lane] = cnt;
            for (int c = 0; c < nch; c++) stg.iatom[co + c] = (unsigned char)lane;
            const int nitems = lp::wave_sum_i(nch);
            if (lane == 0) stg.misc[2] = nitems;
            block_sync<NW>();
            nrebuild++;
            return true;
        };
        if constexpr (PERSIST) {
            // the launch's first exchange: positions as they are (a reset
            // decision moves nothing), lists built from them
            const LayerState L0 = *A.layer;
            fs.dt = L0.dt;
            fs.alpha = L0.alpha;
            fs.fmax = L0.fmax;
            fs.npos = L0.npos;
            fs.steps = L0.steps;
            fs.status = L0.status;
            fs.a = fs.b = 0.f;
            fs.reset = 1;
            fs.dt_move = 0.f;
            msph = own_sphere();
            const float3 z3 = make_float3(0.f, 0.f, 0.f);
            publish(1, z3, z3, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, msph);
            lp_failed = !lp::barrier(LP, ++gen, lane) || !build(1, 0.f);
        }
        for (;;) {
            if (PERSIST && lp_failed) break;
            // parameters of each thread's first angle, dihedral and inversion
            // (read through the caches), issued before the bonds so that their
            // latency overlaps phase A; each term loop prefetches the next
            const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
            float4 apar = tid < td.nangles ? T.ang_par[td.ang_off + tid] : zero4;
            float4 dpar = tid < td.ndih ? T.dih_par[td.dih_off + tid] : zero4;
            float4 ipar = tid < td.ninv ? T.inv_par[td.inv_off + tid] : zero4;
            uint2 idst = tid < td.ninv ? t_invd[tid] : make_uint2(0u, 0u);
            // ---------------- phase A: bonds + hneigh ----------------
            if (!ROT && A.pair_mode)
                for (int w = 0; w < NWP; w++)
                    for (int q = tid; q < N; q += BLOCK) s_fp[w * A.lds_atoms + q] = make_float4(0.f, 0.f, 0.f, 0.f);
            float3 fb[APT];
#pragma unroll
            for (int k = 0; k < APT; k++) {
                const int i = tid + k * BLOCK;
                fb[k] = make_float3(0.f, 0.f, 0.f);
                if (i < N) {
                    const float3 pi = f3(s_pos[i]);
                    const int ngs[4] = {my_ng[k].x, my_ng[k].y, my_ng[k].z, my_ng[k].w};
                    const float bks[4] = {my_bk[k].x, my_bk[k].y, my_bk[k].z, my_bk[k].w};
                    const float bls[4] = {my_bl[k].x, my_bl[k].y, my_bl[k].z, my_bl[k].w};
#pragma unroll
                    for (int sl = 0; sl < 4; sl++) {
                        if (ngs[sl] < 0) {
                            s_h[sl * N + i] = make_float4(0.f, 0.f, 0.f, 1e6f);
                            continue;
                        }
                        float4 hn;
                        (void)bond_term(pi, f3(s_pos[ngs[sl]]), bks[sl], bls[sl], hn, fb[k]);
                        s_h[sl * N + i] = hn;
                    }
                }
            }
            block_sync<NW>();
            // ---------------- phase B: angles, dihedrals, inversions ----------------
            // each force piece goes straight to its atom-major position (the
            // order in which the atom's gather adds it)
            {
                auto put = [&](unsigned p, float3 f) { fint_x[p] = f.x; fint_y[p] = f.y; fint_z[p] = f.z; };
                for (int a = tid; a < td.nangles; a += BLOCK) {
                    const unsigned g = t_ang[a];
                    const uint2 dd = t_angd[a];
                    const float4 par = apar;
                    if (a + BLOCK < td.nangles) apar = T.ang_par[td.ang_off + a + BLOCK];
                    float3 fpi, fpj, fpk;
                    (void)angle_term(s_h[g & 0xffffu], s_h[g >> 16], par, fpi, fpj, fpk);
                    put(dd.x & 0xffffu, fpi); put(dd.x >> 16, fpj); put(dd.y & 0xffffu, fpk);
                }
                for (int d = tid; d < td.ndih; d += BLOCK) {
                    const uint2 g = t_dih[d];
                    const uint2 dd = t_dihd[d];
                    const float4 par = dpar;
                    if (d + BLOCK < td.ndih) dpar = T.dih_par[td.dih_off + d + BLOCK];
                    float3 f1, f2, f3v, f4;
                    (void)dihedral_term(s_h[g.x & 0xffffu], s_h[g.x >> 16], s_h[g.y & 0xffffu], par, (int)(g.y >> 16), f1,
                                        f2, f3v, f4);
                    put(dd.x & 0xffffu, f1); put(dd.x >> 16, f2); put(dd.y & 0xffffu, f3v); put(dd.y >> 16, f4);
                }
                for (int v = tid; v < td.ninv; v += BLOCK) {
                    const uint2 g = t_inv[v];
                    const uint2 dd = idst;
                    const float4 par = ipar;
                    if (v + BLOCK < td.ninv) {
                        ipar = T.inv_par[td.inv_off + v + BLOCK];
                        idst = t_invd[v + BLOCK];
                    }
                    float3 f1, f2, f3v, f4;
                    (void)inversion_term(s_h[g.x & 0xffffu], s_h[g.x >> 16], s_h[g.y], par, f1, f2, f3v, f4);
                    put(dd.x & 0xffffu, f1); put(dd.x >> 16, f2); put(dd.y & 0xffffu, f3v); put(dd.y >> 16, f4);
                }
            }
            block_sync<NW>();
            // ---------------- phase C0: pairs, each unordered pair once ----------------
            // On sweep d the thread of atom i pairs it with j = (i+d) mod N.
            // For fixed d this map is a permutation, so within one wave (all
            // lanes on the same d and k) the reaction force -f_ij goes to a
            // distinct slot of that wave's own LDS buffer: no atomics and a
            // fixed summation order.  Sweeps d = 1..floor(N/2) visit every
            // unordered pair once (for even N the last sweep uses i < N/2).
            float3 pf[APT];
#pragma unroll
            for (int k = 0; k < APT; k++) pf[k] = make_float3(0.f, 0.f, 0.f);
            float3 rr = make_float3(0.f, 0.f, 0.f);   // ROT: reaction forces on the own atom
            if (ROT && A.pair_mode && T.use_pairs) {
                // The accumulators of the reaction sums rotate by one lane per
                // round (DPP): on round t lane v meets row v + t, and the
                // accumulator of row x, which starts in lane x - 1, is then in
                // lane v = x - t.
                //  - up to kSchedMaxAtoms atoms: the lane schedule
                //    (sweep_schedules.h).  Lanes and rows below N are the
                //    atoms, the rest copies, placed so that every pair of
                //    atoms meets within T rounds (PTCDA, 38 atoms: 13 rounds
                //    over all 64 lanes where rows 0..N/2 repeated after the
                //    atoms need 19 with 26 lanes idle).  After the T-1
                //    rotations row x's accumulator is in lane x - T; an atom's
                //    pair force is its lane's own sum plus its rows'
                //    accumulators and its copies' own sums (the schedule's
                //    source list, in a fixed order).
                //  - more atoms: the rotation wraps at N (wave_rot1) over
                //    rounds d = 1..N/2, and lane (i - N/2) mod N ends with atom
                //    i's sum.
                // The loop body is branch-free: the pair force is computed in
                // every lane and zeroed where bit t-1 of `skip` is set (no
                // atom, another meeting sums the pair, a 1-2 neighbour, or the
                // upper half of an even N's last round).
                const bool sch = N <= kSchedMaxAtoms;
                const int nd = N >> 1;
                const int nr = sch ? s_T : nd;
                const bool own = tid < N;
                const int av = sch ? s_av : (own ? tid : -1);
                const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
                const float4 pi4 = av >= 0 ? s_pos[av] : zero;
                // the lane's REQ row with the pair force's constant factors
                // folded in: E0 = -12 sqrtEi sqrtEj and kQ = -qi qj
                const float4 ri0 = av >= 0 ? s_req[av] : zero;
                const float4 ri = make_float4(ri0.x, -12.f * ri0.y, -ri0.z, 0.f);
                unsigned skip = s_skip;
                if (!sch) {
                    const bool even = (N & 1) == 0;
                    skip = own ? 0u : ~0u;
                    const int ngs[4] = {my_ng[0].x, my_ng[0].y, my_ng[0].z, my_ng[0].w};
#pragma unroll
                    for (int sl = 0; sl < 4; sl++) {
                        int dd = ngs[sl] - tid;
                        if (dd < 0) dd += N;
                        if (ngs[sl] >= 0 && dd >= 1 && dd <= nd) skip |= 1u << (dd - 1);
                    }
                    if (even && tid >= nd && nd >= 1) skip |= 1u << (nd - 1);
                }
                KahanF3 pacc;
                pacc.init(make_float3(0.f, 0.f, 0.f));
                // partner of round d: row base + d (the rows after N-1 hold
                // copies: no wrap)
                const int base = sch ? lane : (own ? tid : 0);
                const float4 *prow = s_pos + base;
                const float4 *rrow = s_req + base;
                // one round: the pair's radial factor (zero where skipped)
                // scales dp into the own sum and, negated, into the rotating
                // reaction sum
                auto pair = [&](int d, float4 pc, float4 rc, float3 &bs) {
                    const float3 dp = f3(pc) - f3(pi4);
                    float fr = ljq_pair_fr_folded(dp, ri, rc, T.r2damp);
                    if ((skip >> (d - 1)) & 1u) fr = 0.f;
                    bs.x = __builtin_fmaf(dp.x, fr, bs.x);
                    bs.y = __builtin_fmaf(dp.y, fr, bs.y);
                    bs.z = __builtin_fmaf(dp.z, fr, bs.z);
                    rr.x = __builtin_fmaf(-dp.x, fr, rr.x);
                    rr.y = __builtin_fmaf(-dp.y, fr, rr.y);
                    rr.z = __builtin_fmaf(-dp.z, fr, rr.z);
                };
                // Blocks of four rounds: the four partner rows are read first,
                // then the pairs; each block's own-atom sum goes into the
                // compensated sum.  The last 1-3 rounds form a block of their own.
                auto sweep = [&](auto rotate) {
                    int d0 = 1;
                    for (; d0 + 3 <= nr; d0 += 4) {
                        float4 pc[4], rc[4];
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            pc[u] = prow[d0 + u];
                            rc[u] = rrow[d0 + u];
                        }
                        float3 bs = make_float3(0.f, 0.f, 0.f);
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            if (d0 + u > 1) rotate();
                            pair(d0 + u, pc[u], rc[u], bs);
                        }
                        pacc.add(bs);
                    }
                    if (d0 <= nr) {
                        float3 bs = make_float3(0.f, 0.f, 0.f);
                        for (; d0 <= nr; d0++) {
                            const float4 pc = prow[d0], rc = rrow[d0];
                            if (d0 > 1) rotate();
                            pair(d0, pc, rc, bs);
                        }
                        pacc.add(bs);
                    }
                };
                if (sch) {
                    sweep([&]() {
                        rr.x = wave_rol1(rr.x);
                        rr.y = wave_rol1(rr.y);
                        rr.z = wave_rol1(rr.z);
                    });
                    // the atom's further sums: one byte each (lane << 1 | kind;
                    // kind 1: that lane's accumulator, 0: its own sum), 0xff
                    // ends the list
                    const float3 pv = pacc.value();
                    const unsigned sw = s_src;
                    float3 ext = make_float3(0.f, 0.f, 0.f);
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const unsigned b = (sw >> (8 * k)) & 0xffu;
                        if (__builtin_amdgcn_ballot_w64(b != 0xffu) == 0ull) break;
                        const int sl = (int)(b >> 1) & 63;
                        const float3 xo = make_float3(__shfl(pv.x, sl, 64), __shfl(pv.y, sl, 64), __shfl(pv.z, sl, 64));
                        const float3 xa = make_float3(__shfl(rr.x, sl, 64), __shfl(rr.y, sl, 64), __shfl(rr.z, sl, 64));
                        const float3 x = (b & 1u) ? xa : xo;
                        if (b != 0xffu) ext = ext + x;
                    }
                    pf[0] = pv;
                    rr = ext;
                } else {
                    sweep([&]() {
                        rr.x = wave_rot1(rr.x, N - 1, lane);
                        rr.y = wave_rot1(rr.y, N - 1, lane);
                        rr.z = wave_rot1(rr.z, N - 1, lane);
                    });
                    int src = tid - nd;
                    if (src < 0) src += N;
                    rr.x = __shfl(rr.x, src, 64);
                    rr.y = __shfl(rr.y, src, 64);
                    rr.z = __shfl(rr.z, src, 64);
                    pf[0] = pacc.value();
                }
            } else if (!ROT && A.pair_mode && T.use_pairs) {
                // Several waves: chunk c holds atoms 64c..64c+63, one per
                // lane, and on sweep d = 1..N/2 atom i meets (i+d) mod N.
                // The (chunk, sweep) items, q = c*(N/2) + d-1, are shared
                // evenly by the block's waves in contiguous runs (one or a
                // few chunks each), so that no wave idles while another
                // sweeps a second chunk.  Each wave sums into its own LDS
                // buffer: the reaction -f_ij into slot j as it comes, and a
                // run's own-atom forces (compensated) into slot i at the
                // run's end; the gather adds the waves' buffers.  The
                // neighbour lists of the bonded-pair test are staged in the
                // hneigh area, idle until the next step.
                int4 *s_ngl = reinterpret_cast<int4 *>(s_h);
#pragma unroll
                for (int k = 0; k < APT; k++) {
                    const int i = tid + k * BLOCK;
                    if (i < N) s_ngl[i] = my_ng[k];
                }
                block_sync<NW>();
                float4 *fpw = s_fp + wid * A.lds_atoms;
                const int nd = N >> 1;
                const bool even = (N & 1) == 0;
                const int total = ((N + 63) >> 6) * nd;
                const int share = (total + NW - 1) / NW;
                int q = min(wid * share, total);
                const int qend = min(q + share, total);
                while (q < qend) {   // wave-uniform
                    const int c = q / nd;
                    const int d_lo = q - c * nd + 1;
                    const int d_hi = min(nd, d_lo + (qend - q) - 1);
                    q += d_hi - d_lo + 1;
                    const int i = c * 64 + lane;
                    const bool own = i < N;
                    const float4 pi4 = own ? s_pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
                    // the REQ row with the radial factor's constants folded in
                    // (-12 on sqrtE, the Coulomb sign)
                    const float4 ri0 = own ? s_req[i] : make_float4(0.f, 0.f, 0.f, 0.f);
                    const float4 ri = make_float4(ri0.x, -12.f * ri0.y, -ri0.z, 0.f);
                    const int4 ng = own ? s_ngl[i] : make_int4(-1, -1, -1, -1);
                    // the sweeps this lane drops: its bonded partners on its
                    // own side (sweep (ng - i) mod N), the upper half of an
                    // even N's last sweep, and every sweep of a lane without
                    // an atom; a 64-bit mask per window of 64 sweeps
                    const int ngs[4] = {ng.x, ng.y, ng.z, ng.w};
                    int doff[5];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        int o = ngs[k] - i;
                        if (o < 0) o += N;
                        doff[k] = ngs[k] >= 0 ? o : -1;
                    }
                    doff[4] = (even && i >= nd) ? nd : -1;
                    KahanF3 pacc;
                    pacc.init(make_float3(0.f, 0.f, 0.f));
                    // Positions and REQ rows do not change during the sweep:
                    // the next partner's rows are loaded one sweep ahead (the
                    // index wraps inside the N rows, so the read is in
                    // bounds).  Within a sweep the partners are distinct, so
                    // the reaction slot is read at the top of the iteration
                    // (after the previous sweep's store) and its latency
                    // overlaps the pair.  Branch-free: the mask zeroes the
                    // radial factor of dropped sweeps.  The own forces of four
                    // sweeps are added in FP32 before the compensated add.
                    int jn = (own ? i : 0) + d_lo;   // first partner (d_lo < N)
                    if (jn >= N) jn -= N;
                    float4 pj = s_pos[jn], rj = s_req[jn];
                    for (int w0 = d_lo; w0 <= d_hi; w0 += 64) {
                        const int w1 = min(d_hi, w0 + 63);
                        unsigned long long m = own ? 0ull : ~0ull;
#pragma unroll
                        for (int k = 0; k < 5; k++) {
                            const unsigned t = (unsigned)(doff[k] - w0);
                            if (t < 64u) m |= 1ull << t;
                        }
                        for (int d0 = w0; d0 <= w1; d0 += 4) {
                            const unsigned mb = (unsigned)(m >> (d0 - w0));
                            float3 bs = make_float3(0.f, 0.f, 0.f);
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                if (d0 + u > w1) break;
                                const int j = jn;
                                const float4 pc = pj, rc = rj;
                                jn = (j + 1 == N) ? 0 : j + 1;
                                pj = s_pos[jn];
                                rj = s_req[jn];
                                float4 fj = fpw[j];
                                const float3 dp = f3(pc) - f3(pi4);
                                float fr = ljq_pair_fr_folded(dp, ri, rc, T.r2damp);
                                if ((mb >> u) & 1u) fr = 0.f;
                                const float3 f = dp * fr;
                                bs = bs + f;
                                if (own) {
                                    fj.x -= f.x; fj.y -= f.y; fj.z -= f.z;
                                    fpw[j] = fj;
                                }
                            }
                            pacc.add(bs);
                        }
                    }
                    if (own) {
                        const float3 v = pacc.value();
                        float4 fi = fpw[i];
                        fi.x += v.x; fi.y += v.y; fi.z += v.z;
                        fpw[i] = fi;
                    }
                }
                block_sync<NW>();
            }
            // ---------------- phase C: gather pieces + pairs + surface ----------------
            // surface force of the own atom from its register stencil (one
            // atom per thread; the whole wave takes part in the refresh)
            float4 fe_r = make_float4(0.f, 0.f, 0.f, 0.f);
            if constexpr (RSTEN) {
                if (T.grid) {
                    const bool own = tid < N;
                    const float3 pi = own ? f3(s_pos[tid]) : make_float3(0.f, 0.f, 0.f);
                    const float3 ul = make_float3(pi.x * T.ginv_x, pi.y * T.ginv_y, pi.z * T.ginv_z);
                    const GridCell gc = grid_cell(T.gnx, T.gny, T.gnz, sd.iux, sd.iuy, ul);
                    const int key = gc.iz < 0 ? -2 : (gc.ix * T.gny + gc.iy) * T.gnz + gc.iz;
                    if (own && key != rkey && gc.iz < 0) rkey = key;   // outside the grid: zero surface force
                    const bool need = own && key != rkey;
                    // staging: 4 slots per wave in the hneigh area (the host
                    // reserves >= 16 atoms per wave), idle until the next bonds
                    refresh_stencils(T.grid, T.gnx, T.gny, T.gnz, need, gc, key, my_plq[0], rs, rkey, s_h + wid * 64, 4, lane);
                    if (own) fe_r = grid_eval_stencil([&](int e) { return rs[e]; }, gc);
                }
            }
            if constexpr (PERSIST) {
                // coupled layer, one launch: the list entries of the own
                // atoms in chunks of kLpChunk, lane by lane (the wave's lanes
                // stay busy whatever the molecule's size); each chunk's
                // partial force goes to LDS.  Branch-free: a block's four
                // entries are read first (the chunk's last entry stands in
                // past its end, masked), then their rows, then the pairs; an
                // entry holds the staged atom (bits 0-9) and its instance
                // (bits 10-15).  The switch factor is exactly 1 below its
                // start (x clamped to 0); pairs beyond the cutoff (the skin)
                // contribute nothing.
                const CellGrid &C = A.cells;
                const int nitems = stg.misc[2];
                for (int t = lane; t < nitems; t += 64) {
                    const int a = stg.iatom[t];
                    const int e0 = (t - stg.choff[a]) * kLpChunk;
                    const int n = min(kLpChunk, stg.lcnt[a] - e0);
                    const float3 pa = f3(s_pos[a]);
                    const float4 ra = s_req[a];
                    const unsigned short *nl = stg.nl + a;
                    KahanF3 pk;
                    pk.init(make_float3(0.f, 0.f, 0.f));
                    for (int b0 = 0; b0 < n; b0 += 4) {
                        int ent[4];
#pragma unroll
                        for (int u = 0; u < 4; u++) ent[u] = nl[(e0 + min(b0 + u, n - 1)) * A.lds_atoms];
                        float4 pj[4], rj[4], dg[4];
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            pj[u] = stg.pos[ent[u] & 1023];
                            rj[u] = stg.req[ent[u] & 1023];
                            dg[u] = stg.off[ent[u] >> 10];
                        }
                        float3 bs = make_float3(0.f, 0.f, 0.f);
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const float3 dp = (f3(pj[u]) - pa) + f3(dg[u]);
                            const float r2 = dot3(dp, dp);
                            float3 f;
                            const float e = ljq_pair(dp, ra, rj[u], T.r2damp, f);
                            const float ir = frsq(r2), r = r2 * ir, x = fmaxf((r - C.ron) * C.inv_w, 0.f);
                            const float S = 1.f - x * x * (3.f - 2.f * x), dS_r = 6.f * x * (x - 1.f) * C.inv_w * ir;
                            f = f * S + dp * (e * dS_r);
                            const bool use = b0 + u < n && r2 < C.rc2;
                            bs.x += use ? f.x : 0.f;
                            bs.y += use ? f.y : 0.f;
                            bs.z += use ? f.z : 0.f;
                        }
                        pk.add(bs);
                    }
                    const float3 v = pk.value();
                    stg.ipart[t] = make_float4(v.x, v.y, v.z, 0.f);
                }
                block_sync<NW>();
            }
            // ROT: the gather plan's ranges, all 64 lanes (see the plan above);
            // the own range's rounds and then the helper's sum go into the
            // atom's compensated sum
            [[maybe_unused]] float3 g_own = make_float3(0.f, 0.f, 0.f), g_tail = make_float3(0.f, 0.f, 0.f);
            if constexpr (ROT) {
                KahanF3 gk;
                gk.init(make_float3(0.f, 0.f, 0.f));
                for (int q = 0; q < g_cnt; q += 8) {
                    const float *px = fint_x + g_start + q, *py = fint_y + g_start + q, *pz = fint_z + g_start + q;
                    float3 pv[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) pv[u] = make_float3(px[u], py[u], pz[u]);
                    float3 rs8 = pv[0];
                    if constexpr (FINT_LDS) {
                        // entries past the range (the next atom's pieces,
                        // zeroed slots) weighted by 0
#pragma unroll
                        for (int u = 1; u < 8; u++) {
                            const float w = q + u < g_cnt ? 1.f : 0.f;
                            rs8.x = __builtin_fmaf(pv[u].x, w, rs8.x);
                            rs8.y = __builtin_fmaf(pv[u].y, w, rs8.y);
                            rs8.z = __builtin_fmaf(pv[u].z, w, rs8.z);
                        }
                    } else {
#pragma unroll
                        for (int u = 1; u < 8; u++)
                            if (q + u < g_cnt) rs8 = rs8 + pv[u];
                    }
                    gk.add(rs8);
                }
                g_own = gk.value();
                const int hs = g_help >= 0 ? g_help : lane;
                g_tail = make_float3(__shfl(g_own.x, hs, 64), __shfl(g_own.y, hs, 64), __shfl(g_own.z, hs, 64));
            }
            float3 F[APT];
#pragma unroll
            for (int k = 0; k < APT; k++) {
                const int i = tid + k * BLOCK;
                F[k] = make_float3(0.f, 0.f, 0.f);
                if (i < N) {
                    KahanF3 acc;
                    acc.init(fb[k]);
                    if (A.pair_mode && T.use_pairs) {
                        if constexpr (ROT) {
                            acc.add(pf[k]);
                            acc.add(rr);
                        } else {
                            for (int w = 0; w < NW; w++) acc.add(f3(s_fp[w * A.lds_atoms + i]));
                        }
                    }
                    // the atom's pieces are contiguous (atom-major order); eight
                    // at a time from one base, so that their LDS latencies
                    // overlap.  Each round's pieces are added in list order in
                    // FP32 and the round's sum goes into the compensated
                    // per-atom sum (the rounding error of a round is a few ulp
                    // of its pieces; across rounds it stays compensated).
                    // A round reads up
                    // to 7 entries past the atom's last piece: they lie in the
                    // next atom's pieces, the next array or the 8-float tail
                    // that every force-piece area carries (index < 3*nf_max + 8)
                    if constexpr (ROT) {
                        acc.add(g_own);
                        if (g_help >= 0) acc.add(g_tail);
                    } else {
                        const int pb = my_a2f[k].x, cnt = my_a2f[k].y;
                        for (int q = 0; q < cnt; q += 8) {
                            const float *px = fint_x + pb + q, *py = fint_y + pb + q, *pz = fint_z + pb + q;
                            float3 pv[8];
#pragma unroll
                            for (int u = 0; u < 8; u++) pv[u] = make_float3(px[u], py[u], pz[u]);
                            float3 rs8 = pv[0];
#pragma unroll
                            for (int u = 1; u < 8; u++)
                                if (q + u < cnt) rs8 = rs8 + pv[u];
                            acc.add(rs8);
                        }
                    }
                    const float4 pi4 = s_pos[i];
                    const float3 pi = f3(pi4);
                    if (T.use_pairs && !A.pair_mode) {
                        // every ordered pair (i, j), as getNonBond does
                        const float4 ri = s_req[i];
                        const int4 ng = my_ng[k];
                        for (int j = 0; j < N; j++) {
                            if (j == i || j == ng.x || j == ng.y || j == ng.z || j == ng.w) continue;
                            float3 f;
                            (void)ljq_pair(f3(s_pos[j]) - pi, ri, s_req[j], T.r2damp, f);
                            acc.add(f);
                        }
                    }
                    if constexpr (PERSIST) {
                        // coupled layer, one launch: the partners of the
                        // atom's list (staged neighbour atoms), summed by
                        // chunks above
                        // its chunks' partial forces, in chunk order
                        const int c0 = stg.choff[i], nch = (lcnt + kLpChunk - 1) / kLpChunk;
                        for (int c = 0; c < nch; c++) acc.add(f3(stg.ipart[c0 + c]));
                    } else if (LAYER && A.cells.spos) {
                        // coupled layer: intermolecular forces of the
                        // neighbour cells (k_layer_inter), in cell order
                        const CellGrid &C = A.cells;
                        const int slot = C.slot_of[abase + i];
                        // (cells without partners contribute exact zeros,
                        // skipped); three loads in flight (K = 1, 3, 9, 27)
                        const int K = C.nox * C.noy * C.noz;
                        for (int o0 = 0; o0 < K; o0 += 3) {
                            float4 pf4[3];
#pragma unroll
                            for (int u = 0; u < 3; u++)
                                pf4[u] = o0 + u < K ? C.part[(size_t)(o0 + u) * C.natoms + slot]
                                                    : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                            for (int u = 0; u < 3; u++)
                                if (pf4[u].x != 0.f || pf4[u].y != 0.f || pf4[u].z != 0.f) acc.add(f3(pf4[u]));
                        }
                    }
                    if (T.grid) {
                        float3 ul = make_float3(pi.x * T.ginv_x, pi.y * T.ginv_y, pi.z * T.ginv_z);
                        float4 fe;
                        if constexpr (RSTEN) {
                            fe = fe_r;
                        } else if (A.stencil) {
                            GridCell gc = grid_cell(T.gnx, T.gny, T.gnz, sd.iux, sd.iuy, ul);
                            const int key = gc.iz < 0 ? -2 : (gc.ix * T.gny + gc.iy) * T.gnz + gc.iz;
                            // the system's block of 16*N float4, element-major
                            float4 *sc = A.stencil + (size_t)abase * 16 + i;
                            if (key != skey[k]) {
                                if (gc.iz >= 0) grid_store_stencil(T.grid, T.gnx, T.gny, T.gnz, gc, my_plq[k], sc, N);
                                skey[k] = key;
                            }
                            fe = grid_eval_cached(sc, N, gc);
                        } else {
                            fe = grid_sample(T.grid, T.gnx, T.gny, T.gnz, sd.iux, sd.iuy, ul, my_plq[k]);
                        }
                        acc.add(make_float3(-fe.x * T.ginv_x, -fe.y * T.ginv_y, -fe.z * T.ginv_z));
                    }
                    F[k] = acc.value();
                }
            }
            // ---------------- phase D: integrator ----------------
            if (A.mode == KMODE_EVAL) {
#pragma unroll
                for (int k = 0; k < APT; k++) {
                    const int i = tid + k * BLOCK;
                    if (i < N) {
                        A.force[abase + i] = make_float4(F[k].x, F[k].y, F[k].z, 0.f);
                        if (k == 0) persist_stencil();
                        if (A.stencil_key) A.stencil_key[abase + i] = skey[k];
                    }
                }
                double f2m = 0.0;
#pragma unroll
                for (int k = 0; k < APT; k++)
                    if (tid + k * BLOCK < N) {
                        double fx = F[k].x, fy = F[k].y, fz = F[k].z;
                        f2m = fmax(f2m, fx * fx + fy * fy + fz * fz);
                    }
                f2m = wave_max(f2m);
                if (lane == 0) s_red[wid * 8 + 6] = f2m;
                block_sync<NW>();
                if (tid == 0) {
                    double m = 0.0;
                    for (int w = 0; w < NW; w++) m = fmax(m, s_red[w * 8 + 6]);
                    // energies: k_energy64 at the same positions
                    st.fmax = sqrt(m);
                    A.state[s] = st;
                }
                block_sync<NW>();
                break;
            }
            // per-atom partial sums (FP64)
            double p_vf = 0.0, p_vv = 0.0, p_ff = 0.0, p_fm = 0.0;
            if (A.mode == KMODE_FIRE || A.mode == KMODE_LAYER || A.mode == KMODE_LAYER_P) {
                const float dtf = (float)(PERSIST ? fs.dt : (LAYER && A.mode == KMODE_LAYER ? A.layer->dt : st.dt));
                const double flim = A.fire.flim;
#pragma unroll
                for (int k = 0; k < APT; k++) {
                    if (tid + k * BLOCK < N) {
                        double fx = F[k].x, fy = F[k].y, fz = F[k].z;
                        const double f2 = fx * fx + fy * fy + fz * fz;
                        p_fm = fmax(p_fm, f2);
                        // |F| clamp of step_fire (raff.rs:1089-1096): the
                        // clamped force drives v and v.F, |F|^2 sums the
                        // unclamped one
                        p_ff += f2;
                        if (flim > 0.0 && f2 > flim * flim) {
                            float sc = (float)(flim / sqrt(f2));
                            F[k] = F[k] * sc;
                            fx = F[k].x; fy = F[k].y; fz = F[k].z;
                        }
                        my_v[k] = my_v[k] + F[k] * dtf;
                        double vx = my_v[k].x, vy = my_v[k].y, vz = my_v[k].z;
                        p_vf += vx * fx + vy * fy + vz * fz;
                        p_vv += vx * vx + vy * vy + vz * vz;
                    }
                }
            } else {  // KMODE_MD: MolWorld::run_md / move_atom_md
                const float dtf = (float)A.fire.md_dt;
                const float cd = (float)A.fire.md_cdamp;
                const double flim = A.fire.md_flim;
#pragma unroll
                for (int k = 0; k < APT; k++) {
                    const int i = tid + k * BLOCK;
                    if (i < N) {
                        double fx = F[k].x, fy = F[k].y, fz = F[k].z;
                        double vx = my_v[k].x, vy = my_v[k].y, vz = my_v[k].z;
                        double f2 = fx * fx + fy * fy + fz * fz;
                        p_vf += vx * fx + vy * fy + vz * fz;
                        p_ff += f2;
                        p_fm = fmax(p_fm, f2);
                        float3 fc = F[k];
                        if (f2 > flim * flim) fc = fc * (float)(flim / sqrt(f2));
                        my_v[k] = my_v[k] * cd + fc * dtf;
                    }
                }
            }
            if constexpr (PERSIST) {
                // ---- coupled layer, one launch: end of the step ----
                const bool own = tid < N;
                const float3 pcur = own ? f3(s_pos[tid]) : make_float3(0.f, 0.f, 0.f);
                const float3 dpr = pcur - pref;
                const double vm2 = own ? (double)dot3(my_v[0], my_v[0]) : 0.0;
                const double dm2 = own ? (double)dot3(dpr, dpr) : 0.0;
                msph = own_sphere();
                const double r0 = wave_sum(p_vf), r1 = wave_sum(p_vv), r2 = wave_sum(p_ff), r3 = wave_max(p_fm),
                             r4 = wave_max(vm2), r5 = wave_max(dm2);
                const int par = kstep & 1;
                publish(par, my_v[0], F[0], r0, r1, r2, r3, r4, r5, msph);
                if (!lp::barrier(LP, ++gen, lane)) {
                    lp_failed = true;
                    break;
                }
                // every workgroup: the records of all molecules in the same
                // order (k_layer_decide's), the same decision
                double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0, a5 = 0.0;
                for (int q0 = lane; q0 < (int)nmol; q0 += 8 * 64) {
                    float4 x0[8], x1[8], x2[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int q = q0 + 64 * u;
                        if (q < (int)nmol) {
                            const unsigned ro = (((unsigned)par * nmol + q) * kLpRec) * 8u;
                            x0[u] = lp::ldc4(rrec, ro);
                            x1[u] = lp::ldc4(rrec, ro + 16u);
                            x2[u] = lp::ldc4(rrec, ro + 32u);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        if (q0 + 64 * u < (int)nmol) {
                            const double2 d01 = lp::f42d(x0[u]), d23 = lp::f42d(x1[u]), d45 = lp::f42d(x2[u]);
                            a0 += d01.x;
                            a1 += d01.y;
                            a2 += d23.x;
                            a3 = fmax(a3, d23.y);
                            a4 = fmax(a4, d45.x);
                            a5 = fmax(a5, d45.y);
                        }
                    }
                }
                const double vf = wave_sum(a0), vv = wave_sum(a1), ff = wave_sum(a2), fm = wave_max(a3),
                             vm = wave_max(a4), dm = wave_max(a5);
                lp::decide(fs, A.fire, vf, vv, ff, fm);
                // the move of the own atoms
                if (own) {
                    const float3 v2 = lp::mixed_velocity(my_v[0], F[0], fs.a, fs.b, fs.reset);
                    const float3 p2 = lp::moved(pcur, v2, fs.dt_move);
                    my_v[0] = v2;
                    const float4 p4 = make_float4(p2.x, p2.y, p2.z, s_pos[tid].w);
                    s_pos[tid] = p4;
                }
                copy_rows(false);
                kstep++;
                step_in_launch++;
                if (fs.status != SMR_SYS_RUNNING || step_in_launch >= nsteps) {
                    if (own) {
                        persist_stencil();
                        if (A.stencil_key) A.stencil_key[abase + tid] = skey[0];
                        A.pos[abase + tid] = s_pos[tid];
                        A.vel[abase + tid] = make_float4(my_v[0].x, my_v[0].y, my_v[0].z, 0.f);
                    }
                    if (blockIdx.x == 0 && tid == 0) {
                        LayerState S = *A.layer;
                        S.dt = fs.dt;
                        S.alpha = fs.alpha;
                        S.npos = fs.npos;
                        S.steps = fs.steps;
                        S.status = fs.status;
                        S.fmax = fs.fmax;
                        S.pending_move = 0;
                        S.cells_dirty = 1;
                        S.rebuilds += nrebuild;
                        *A.layer = S;
                    }
                    break;
                }
                // the lists hold while no atom has moved more than skin/2
                // since they were built; this step moves an atom by at most
                // dt (a |v| + b |F|)
                const float delta =
                    fs.reset ? 0.f : fs.dt_move * (fs.a * (float)sqrt(vm) + fs.b * (float)sqrt(fm));
                if ((float)sqrt(dm) + delta > 0.5f * LP.skin) {
                    if (!build(par, delta)) {
                        lp_failed = true;
                        break;
                    }
                } else {
                    restage(par);
                }
                block_sync<NW>();
                continue;
            }
            // block reduction of (vf, vv, ff, fmax^2)
            if (LAYER && A.mode == KMODE_LAYER) {
                // coupled layer: clamped forces and updated velocities go to
                // memory, this molecule's sums to its partial slot; the shared
                // FIRE decision is taken by the next launch (k_layer_decide)
#pragma unroll
                for (int k = 0; k < APT; k++) {
                    const int i = tid + k * BLOCK;
                    if (i < N) {
                        A.force[abase + i] = make_float4(F[k].x, F[k].y, F[k].z, 0.f);
                        A.vel[abase + i] = make_float4(my_v[k].x, my_v[k].y, my_v[k].z, 0.f);
                        if (k == 0) persist_stencil();
                        if (A.stencil_key) A.stencil_key[abase + i] = skey[k];
                    }
                }
                double r0 = wave_sum(p_vf), r1 = wave_sum(p_vv), r2 = wave_sum(p_ff), r3 = wave_max(p_fm);
                if (lane == 0) {
                    s_red[wid * 8 + 0] = r0; s_red[wid * 8 + 1] = r1;
                    s_red[wid * 8 + 2] = r2; s_red[wid * 8 + 3] = r3;
                }
                block_sync<NW>();
                if (tid == 0) {
                    r0 = r1 = r2 = r3 = 0.0;
                    for (int w = 0; w < NW; w++) {
                        r0 += s_red[w * 8 + 0]; r1 += s_red[w * 8 + 1];
                        r2 += s_red[w * 8 + 2]; r3 = fmax(r3, s_red[w * 8 + 3]);
                    }
                    A.partials[s] = make_double4(r0, r1, r2, r3);
                }
                block_sync<NW>();
                break;
            }
            // Convergence (max |F|^2 < fconv^2) is tested as "no atom has
            // |F|^2 >= fconv^2" (a ballot); max |F|^2 itself is reduced only
            // when the system leaves the launch.
            bool any_big;
            {
                const double fc2 = A.fire.fconv * A.fire.fconv;
                any_big = __builtin_amdgcn_ballot_w64(p_fm >= fc2) != 0ull;
                double r0, r1, r2;
                wave_sum3(p_vf, p_vv, p_ff, r0, r1, r2);
                if (NW > 1) {
                    if (lane == 0) {
                        s_red[wid * 8 + 0] = r0; s_red[wid * 8 + 1] = r1;
                        s_red[wid * 8 + 2] = r2; s_red[wid * 8 + 3] = any_big ? 1.0 : 0.0;
                    }
                    block_sync<NW>();
                    r0 = r1 = r2 = 0.0;
                    any_big = false;
                    for (int w = 0; w < NW; w++) {
                        r0 += s_red[w * 8 + 0]; r1 += s_red[w * 8 + 1];
                        r2 += s_red[w * 8 + 2]; any_big = any_big || s_red[w * 8 + 3] != 0.0;
                    }
                    block_sync<NW>();
                }
                p_vf = r0; p_vv = r1; p_ff = r2;
            }
            const double dt_old = st.dt;
            float dt_move = (float)dt_old;
            bool nonfinite = !(isfinite(p_vf) && isfinite(p_vv) && isfinite(p_ff));
            FireConst FC;
            {
                ConstArgs *fa = fresh_args();
                FC.dt_max = fa->fire.dt_max; FC.alpha0 = fa->fire.alpha0; FC.f_inc = fa->fire.f_inc;
                FC.f_dec = fa->fire.f_dec; FC.f_alpha = fa->fire.f_alpha; FC.fconv = fa->fire.fconv;
                FC.n_min = fa->fire.n_min; FC.max_steps = fa->fire.max_steps; FC.md_dt = fa->fire.md_dt;
            }
            if (A.mode == KMODE_FIRE) {
                // FIRE decision (relax_pentacene_uff.rs:162-181)
                if (p_vf > 0.0 && p_ff > 1e-30 && p_vv > 1e-30) {
                    // the mixing factors are applied in FP32: alpha |v| / |F|
                    // from the FP64 sums with FP32 roots
                    const float af = (float)(1.0 - st.alpha);
                    const float vvf = (float)p_vv;
                    const float bf = (float)st.alpha * (vvf * frsq(vvf)) * frsq((float)p_ff);
#pragma unroll
                    for (int k = 0; k < APT; k++) my_v[k] = my_v[k] * af + F[k] * bf;
                    st.npos++;
                    if (st.npos > FC.n_min) {
                        st.dt = fmin(st.dt * FC.f_inc, FC.dt_max);
                        st.alpha *= FC.f_alpha;
                    }
                } else {
                    st.npos = 0;
                    st.dt *= FC.f_dec;
                    st.alpha = FC.alpha0;
#pragma unroll
                    for (int k = 0; k < APT; k++) my_v[k] = make_float3(0.f, 0.f, 0.f);
                }
            } else {
                dt_move = (float)FC.md_dt;
                if (p_vf < 0.0) {
                    // v is zeroed after the move (molecular.rs:156-158)
                }
            }
            // move (x += v*dt_old), own atoms only; every lane has finished
            // reading positions for this step's pairs before any is moved
            block_sync<NW>();
#pragma unroll
            for (int k = 0; k < APT; k++) {
                const int i = tid + k * BLOCK;
                if (i < N) {
                    float4 p = s_pos[i];
                    p.x += my_v[k].x * dt_move;
                    p.y += my_v[k].y * dt_move;
                    p.z += my_v[k].z * dt_move;
                    s_pos[i] = p;
                }
            }
            copy_rows(false);
            if (A.mode == KMODE_MD && p_vf < 0.0) {
#pragma unroll
                for (int k = 0; k < APT; k++) my_v[k] = make_float3(0.f, 0.f, 0.f);
            }
            st.steps++;
            step_in_launch++;
            bool done = false;
            if (nonfinite) {
                st.status = SMR_SYS_NONFINITE;
                done = true;
            } else if (A.mode == KMODE_FIRE ? !any_big : (p_ff < FC.fconv * FC.fconv)) {
                st.status = SMR_SYS_CONVERGED;
                done = true;
            } else if (st.steps >= FC.max_steps) {
                st.status = SMR_SYS_MAXSTEPS;
                done = true;
            }
            const bool leave = done || step_in_launch >= nsteps;
            if (leave) {
                // max |F|^2 of this step's forces
                double fm2 = wave_max(p_fm);
                if (NW > 1) {
                    if (lane == 0) s_red[wid * 8 + 4] = fm2;
                    block_sync<NW>();
                    fm2 = 0.0;
                    for (int w = 0; w < NW; w++) fm2 = fmax(fm2, s_red[w * 8 + 4]);
                }
                // the energies of the final positions are evaluated in FP64
                // afterwards (k_energy64)
                if (tid == 0) {
                    st.fmax = sqrt(fm2);
                    fa->state[s] = st;
                }
#pragma unroll
                for (int k = 0; k < APT; k++) {
                    const int i = tid + k * BLOCK;
                    if (i < N) {
                        if (k == 0) persist_stencil();
                        if (fa->stencil_key) fa->stencil_key[abase + i] = skey[k];
                        fa->pos[abase + i] = s_pos[i];
                        fa->vel[abase + i] = make_float4(my_v[k].x, my_v[k].y, my_v[k].z, 0.f);
                    }
                }
                block_sync<NW>();
                break;
            }
            block_sync<NW>();
        }
    }
}

// ===========================================================================
// plain method: one kernel per term, one system
// ===========================================================================
__global__ void k_plain_bonds(PlainArgs P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const TopoDesc td = P.tab.topo[P.topo];
    if (i >= td.natoms) return;
    const int ta = td.atom_off + i;
    const int4 ng = P.tab.ng[ta];
    const float4 bk = P.tab.bk[ta], bl = P.tab.bl[ta];
    const int ngs[4] = {ng.x, ng.y, ng.z, ng.w};
    const float bks[4] = {bk.x, bk.y, bk.z, bk.w}, bls[4] = {bl.x, bl.y, bl.z, bl.w};
    const float3 pi = f3(P.pos[i]);
    float3 f = make_float3(0.f, 0.f, 0.f);
    float eb = 0.f;
    for (int sl = 0; sl < 4; sl++) {
        if (ngs[sl] < 0) { P.hneigh[sl * td.natoms + i] = make_float4(0.f, 0.f, 0.f, 1e6f); continue; }
        float4 hn;
        eb += bond_term(pi, f3(P.pos[ngs[sl]]), bks[sl], bls[sl], hn, f);
        P.hneigh[sl * td.natoms + i] = hn;
    }
    P.force[i] = make_float4(f.x, f.y, f.z, 0.f);
    P.eatom[i] = eb;
}

__global__ void k_plain_angles(PlainArgs P) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const TopoDesc td = P.tab.topo[P.topo];
    if (a >= td.nangles) return;
    const unsigned g = P.tab.ang_ng[td.ang_off + a];
    float3 fpi, fpj, fpk;
    float e = angle_term(P.hneigh[g & 0xffffu], P.hneigh[g >> 16], P.tab.ang_par[td.ang_off + a], fpi, fpj, fpk);
    const int p = 4 * td.ndih + 4 * td.ninv + a, q = td.nangles;
    P.fint[p] = make_float4(fpi.x, fpi.y, fpi.z, e);
    P.fint[p + q] = make_float4(fpj.x, fpj.y, fpj.z, 0.f);
    P.fint[p + 2 * q] = make_float4(fpk.x, fpk.y, fpk.z, 0.f);
}

__global__ void k_plain_dihedrals(PlainArgs P) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const TopoDesc td = P.tab.topo[P.topo];
    if (d >= td.ndih) return;
    const uint2 g = P.tab.dih_ng[td.dih_off + d];
    float3 f1, f2, f3v, f4;
    float e = dihedral_term(P.hneigh[g.x & 0xffffu], P.hneigh[g.x >> 16], P.hneigh[g.y & 0xffffu],
                            P.tab.dih_par[td.dih_off + d], (int)(g.y >> 16), f1, f2, f3v, f4);
    const int p = d, q = td.ndih;
    P.fint[p] = make_float4(f1.x, f1.y, f1.z, e);
    P.fint[p + q] = make_float4(f2.x, f2.y, f2.z, 0.f);
    P.fint[p + 2 * q] = make_float4(f3v.x, f3v.y, f3v.z, 0.f);
    P.fint[p + 3 * q] = make_float4(f4.x, f4.y, f4.z, 0.f);
}

__global__ void k_plain_inversions(PlainArgs P) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const TopoDesc td = P.tab.topo[P.topo];
    if (v >= td.ninv) return;
    const uint2 g = P.tab.inv_ng[td.inv_off + v];
    float3 f1, f2, f3v, f4;
    float e = inversion_term(P.hneigh[g.x & 0xffffu], P.hneigh[g.x >> 16], P.hneigh[g.y], P.tab.inv_par[td.inv_off + v],
                             f1, f2, f3v, f4);
    const int p = 4 * td.ndih + v, q = td.ninv;
    P.fint[p] = make_float4(f1.x, f1.y, f1.z, e);
    P.fint[p + q] = make_float4(f2.x, f2.y, f2.z, 0.f);
    P.fint[p + 2 * q] = make_float4(f3v.x, f3v.y, f3v.z, 0.f);
    P.fint[p + 3 * q] = make_float4(f4.x, f4.y, f4.z, 0.f);
}

// assemble + pairs + surface (assembleForces_UFF + getNonBond_GridFF_Bspline)
__global__ void k_plain_assemble(PlainArgs P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const TopoDesc td = P.tab.topo[P.topo];
    if (i >= td.natoms) return;
    const int ta = td.atom_off + i;
    KahanF3 acc;
    acc.init(f3(P.force[i]));
    const int2 se = P.tab.a2f[ta];
    const unsigned short *idx = P.tab.a2f_idx + td.a2f_off + se.x;
    for (int q = 0; q < se.y; q++) acc.add(f3(P.fint[idx[q]]));
    const float3 pi = f3(P.pos[i]);
    float enb = 0.f, esurf = 0.f;
    double enb_d = 0.0;
    if (P.tab.use_pairs) {
        const float4 ri = P.tab.req[ta];
        const int4 ng = P.tab.ng[ta];
        for (int j = 0; j < td.natoms; j++) {
            if (j == i || j == ng.x || j == ng.y || j == ng.z || j == ng.w) continue;
            float3 f;
            float e = ljq_pair(f3(P.pos[j]) - pi, ri, P.tab.req[td.atom_off + j], P.tab.r2damp, f);
            acc.add(f);
            enb_d += (double)e;
        }
    }
    (void)enb;
    if (P.tab.grid) {
        float3 ul = make_float3(pi.x * P.tab.ginv_x, pi.y * P.tab.ginv_y, pi.z * P.tab.ginv_z);
        float4 fe = grid_sample(P.tab.grid, P.tab.gnx, P.tab.gny, P.tab.gnz, P.iux, P.iuy, ul, P.tab.plq[ta]);
        acc.add(make_float3(-fe.x * P.tab.ginv_x, -fe.y * P.tab.ginv_y, -fe.z * P.tab.ginv_z));
        esurf = fe.w;
    }
    float3 F = acc.value();
    P.force[i] = make_float4(F.x, F.y, F.z, 0.f);
    P.epair[i] = enb_d;
    P.esurf[i] = esurf;
}

// per-system FIRE reductions: v += F dt, (v.F, |v|^2, |F|^2, max|F|^2) and
// energy terms, in one workgroup (FP64).
__global__ void k_plain_fire_reduce(PlainArgs P, double dt, double flim) {
    const TopoDesc td = P.tab.topo[P.topo];
    __shared__ double red[16][11];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
    double r[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const float dtf = (float)dt;
    for (int i = tid; i < td.natoms; i += blockDim.x) {
        float4 F4 = P.force[i];
        float3 F = f3(F4);
        double fx = F.x, fy = F.y, fz = F.z;
        const double f2 = fx * fx + fy * fy + fz * fz;
        r[3] = fmax(r[3], f2);
        r[2] += f2;   // unclamped |F|^2, as step_fire (raff.rs:1089-1096)
        if (flim > 0.0 && f2 > flim * flim) {
            F = F * (float)(flim / sqrt(f2));
            P.force[i] = make_float4(F.x, F.y, F.z, 0.f);
            fx = F.x; fy = F.y; fz = F.z;
        }
        float4 v4 = P.vel[i];
        float3 v = f3(v4) + F * dtf;
        P.vel[i] = make_float4(v.x, v.y, v.z, 0.f);
        double vx = v.x, vy = v.y, vz = v.z;
        r[0] += vx * fx + vy * fy + vz * fz;
        r[1] += vx * vx + vy * vy + vz * vz;
        r[4] += (double)P.eatom[i];
        r[8] += P.epair[i];
        r[9] += (double)P.esurf[i];
    }
    for (int a = tid; a < td.nangles; a += blockDim.x) r[5] += (double)P.fint[4 * td.ndih + 4 * td.ninv + a].w;
    for (int d = tid; d < td.ndih; d += blockDim.x) r[6] += (double)P.fint[d].w;
    for (int v = tid; v < td.ninv; v += blockDim.x) r[7] += (double)P.fint[4 * td.ndih + v].w;
    for (int q = 0; q < 11; q++) r[q] = (q == 3) ? wave_max(r[q]) : wave_sum(r[q]);
    if (lane == 0)
        for (int q = 0; q < 11; q++) red[wid][q] = r[q];
    __syncthreads();
    if (tid == 0) {
        double o[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int w = 0; w < nw; w++)
            for (int q = 0; q < 11; q++) o[q] = (q == 3) ? fmax(o[q], red[w][q]) : o[q] + red[w][q];
        // out: vf, vv, ff, f2max, Eb, Ea, Ed, Ei, Enb, Esurf
        for (int q = 0; q < 10; q++) P.cvf[q] = o[q];
    }
}

// FIRE velocity mixing / reset and move, coefficients decided on the host.
__global__ void k_plain_fire_apply(PlainArgs P, float a, float b, int reset, float dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const TopoDesc td = P.tab.topo[P.topo];
    if (i >= td.natoms) return;
    float3 v = f3(P.vel[i]);
    if (reset) v = make_float3(0.f, 0.f, 0.f);
    else v = v * a + f3(P.force[i]) * b;
    P.vel[i] = make_float4(v.x, v.y, v.z, 0.f);
    float4 p = P.pos[i];
    p.x += v.x * dt;
    p.y += v.y * dt;
    p.z += v.z * dt;
    P.pos[i] = p;
}

// ===========================================================================
// host-side launch helpers
// ===========================================================================
template <int BLOCK, int APT, bool FL>
static hipError_t launch_one(const KernelArgs &A, int grid, size_t lds, hipStream_t stream) {
    if (A.cells.spos)
        hipLaunchKernelGGL((k_relax<BLOCK, APT, FL, true>), dim3(grid), dim3(BLOCK), lds, stream, A);
    else
        hipLaunchKernelGGL((k_relax<BLOCK, APT, FL, false>), dim3(grid), dim3(BLOCK), lds, stream, A);
    return hipGetLastError();
}

hipError_t launch_relax_layer_persistent(const KernelArgs &A, int grid, size_t lds, hipStream_t stream) {
    void *args[] = {const_cast<KernelArgs *>(&A)};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_relax<64, 1, true, true, true>), dim3(grid),
                                      dim3(64), args, (unsigned)lds, stream);
}

const void *relax_layer_persistent_ptr() { return reinterpret_cast<const void *>(&k_relax<64, 1, true, true, true>); }

// (occupancy of a shape: the layer instances use the same launch bounds)
template <int BLOCK, int APT, bool FL>
static const void *kernel_ptr() {
    return reinterpret_cast<const void *>(&k_relax<BLOCK, APT, FL, false>);
}

// instantiated shapes: (threads per workgroup, atoms per thread)
#define SMR_SHAPES(X) \
    X(64, 1) X(64, 2) X(64, 4) X(128, 1) X(128, 2) X(128, 4) X(256, 1) X(256, 2) X(256, 4) X(512, 1) X(512, 2)

hipError_t launch_relax(const KernelArgs &A, const LaunchShape &L, hipStream_t stream) {
#define SMR_CASE(B, P)                                                                                   \
    if (L.block == B && L.apt == P)                                                                      \
        return L.fint_lds ? launch_one<B, P, true>(A, L.grid, L.lds_bytes, stream)                       \
                          : launch_one<B, P, false>(A, L.grid, L.lds_bytes, stream);
    SMR_SHAPES(SMR_CASE)
#undef SMR_CASE
    return hipErrorInvalidValue;
}

const void *relax_kernel_ptr(int block, int apt, bool fint_lds) {
#define SMR_CASE(B, P) \
    if (block == B && apt == P) return fint_lds ? kernel_ptr<B, P, true>() : kernel_ptr<B, P, false>();
    SMR_SHAPES(SMR_CASE)
#undef SMR_CASE
    return nullptr;
}

// zero the per-class work counters (a kernel, not a memset: the launch
// sequence may be captured into a graph)
__global__ void k_zero_i32(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_zero_i32(int *p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_plain_step_forces(const PlainArgs &P, int natoms, int nang, int ndih, int ninv, hipStream_t st) {
    const int B = 64;
    auto nb = [](int n) { return n > 0 ? (n + 63) / 64 : 1; };
    hipLaunchKernelGGL(k_plain_bonds, dim3(nb(natoms)), dim3(B), 0, st, P);
    if (nang > 0) hipLaunchKernelGGL(k_plain_angles, dim3(nb(nang)), dim3(B), 0, st, P);
    if (ndih > 0) hipLaunchKernelGGL(k_plain_dihedrals, dim3(nb(ndih)), dim3(B), 0, st, P);
    if (ninv > 0) hipLaunchKernelGGL(k_plain_inversions, dim3(nb(ninv)), dim3(B), 0, st, P);
    hipLaunchKernelGGL(k_plain_assemble, dim3(nb(natoms)), dim3(B), 0, st, P);
    return hipGetLastError();
}

hipError_t launch_plain_reduce(const PlainArgs &P, double dt, double flim, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_fire_reduce, dim3(1), dim3(256), 0, st, P, dt, flim);
    return hipGetLastError();
}

hipError_t launch_plain_apply(const PlainArgs &P, int natoms, float a, float b, int reset, float dt, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_fire_apply, dim3((natoms + 63) / 64), dim3(64), 0, st, P, a, b, reset, dt);
    return hipGetLastError();
}

}  // namespace smr
