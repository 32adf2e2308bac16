// raff_kernels.hip -- multi-replica rigid-atom port relaxation (RAFF "RRsp3"
// cluster step) on gfx950.
//
// Reference: oclff::RRsp3Multi::step_cluster_multi (crates/libs/oclff/src/
// rrsp3.rs:1028-1050) with the kernels compute_collision_multi,
// compute_ports_current_multi and apply_corrections_multi
// (opencl/RRsp3.cl:1774-2047).  One reference step is three launches over a
// (group, replica) grid, exchanging every correction through global memory.
// Here one persistent launch runs many steps: each replica lives in a
// segment of SEG lanes (one atom per lane; two replicas share a wave64 when
// they have at most 32 atoms), its positions sit in LDS for the pair and
// port loops, the port recoils pass through an LDS slot array, and every
// per-atom state (position, orientation, heavy-ball momenta, shared topology
// rows and port arms) stays in registers across the steps.  Global memory is
// read at the start of a launch and written at its end.
//
// The arithmetic of a step follows the three reference kernels in order:
//   collision  Jacobi sphere-pair corrections, 1-2 and 1-3 pairs excluded
//   ports      per node and port k: tip = x_i + R(q_i) r_k, constraint
//              C = |x_j - tip|, XPBD weights w_i + w_j + |r x n|^2 / I + 1/(K dt^2)
//              (the angular weight twice and the impulse halved when j is a
//              node: both ends correct the bond), node translation and
//              rotation sums, the recoil of j into its port slot
//   apply      dx = dx_coll + dx_node + recoils gathered through bk_slots;
//              fix mask; x += relaxation dx + beta dx_prev; nodes rotate by
//              relaxation dtheta, heavy-ball quaternion momentum, sign kept
//              continuous.
#include <hip/hip_runtime.h>

#include "smr_kernels.h"

namespace smr {

namespace {

__device__ __forceinline__ float3 v3(float4 a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float3 vadd(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 vsub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 vmul(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float vdot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 vcross(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// v rotated by the unit quaternion q = (x, y, z, w):  v + w t + q_v x t,
// t = 2 q_v x v
__device__ __forceinline__ float3 qrot(float4 q, float3 v) {
    const float3 u = make_float3(q.x, q.y, q.z);
    const float3 t = vmul(vcross(u, v), 2.0f);
    return vadd(vadd(v, vmul(t, q.w)), vcross(u, t));
}

// Hamilton product a * b
__device__ __forceinline__ float4 qmul(float4 a, float4 b) {
    return make_float4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                       a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                       a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
                       a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}

__device__ __forceinline__ float4 qnormalize(float4 q) {
    const float n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (n2 < 1e-16f) return make_float4(0.f, 0.f, 0.f, 1.f);
    const float s = rsqrtf(n2);
    return make_float4(q.x * s, q.y * s, q.z * s, q.w * s);
}

// rotation by `angle` about `axis` (normalised here)
__device__ __forceinline__ float4 qaxis(float3 axis, float angle) {
    const float a = sqrtf(vdot(axis, axis));
    if (a < 1e-8f || fabsf(angle) < 1e-8f) return make_float4(0.f, 0.f, 0.f, 1.f);
    const float3 u = vmul(axis, 1.0f / a);
    float s, c;
    sincosf(0.5f * angle, &s, &c);
    return make_float4(u.x * s, u.y * s, u.z * s, c);
}

__device__ __forceinline__ bool excluded(int j, int4 a, int4 b) {
    return j == a.x || j == a.y || j == a.z || j == a.w || j == b.x || j == b.y || j == b.z || j == b.w;
}

// max over the SEG lanes of a segment (SEG <= 64: inside one wave)
template <int SEG>
__device__ __forceinline__ float seg_max(float v) {
#pragma unroll
    for (int off = SEG / 2; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, SEG));
    return v;
}

}  // namespace

// One replica per segment of SEG lanes; BLK = max(SEG, 64) threads.
template <int SEG>
__global__ __launch_bounds__(SEG < 64 ? 64 : SEG) __attribute__((amdgpu_waves_per_eu(4))) void k_raff(RaffArgs A) {
    constexpr int BLK = SEG < 64 ? 64 : SEG;
    constexpr int R = BLK / SEG;            // replicas per workgroup
    constexpr int NW = BLK / 64;
    // one replica per wave at most, and no more ports than lanes: the ports
    // are spread over the lanes (one lane per port) instead of looping over
    // each node's four slots
    constexpr bool SPREAD = SEG <= 64;
    constexpr int PB = SPREAD ? BLK : 1;
    __shared__ float4 l_pos[BLK];
    __shared__ float4 l_slot[BLK * 4];      // port recoils: [replica][node*4 + k]
    __shared__ float l_rad[BLK];
    __shared__ float l_red[R * NW];
    __shared__ int l_run[R];
    __shared__ float4 l_quat[PB];           // node orientations of the step
    __shared__ float4 l_pdp[PB], l_pdt[PB]; // per port: node translation, rotation contributions
    __shared__ int l_port[PB];              // per replica: its ports (node*4 + k | neighbour << 16), node-major
    __shared__ float4 l_parm[PB];           // ... their arms (xyz) and stiffnesses (w)
    __shared__ int l_pfirst[PB];            // per node: first port | port count << 16
    __shared__ int l_nport[R];

    const int tid = threadIdx.x;
    const int r = tid / SEG, lid = tid % SEG;
    const int sys = blockIdx.x * R + r;
    const int n = A.natoms, nnode = A.nnode;
    const bool inrange = sys < A.nsys && lid < n;
    const size_t ia = (size_t)sys * n + lid;

    float4 p4 = make_float4(0.f, 0.f, 0.f, 0.f), q = make_float4(0.f, 0.f, 0.f, 1.f);
    float4 mom = make_float4(0.f, 0.f, 0.f, 0.f), qmom = make_float4(0.f, 0.f, 0.f, 0.f);
    int4 ng = make_int4(-1, -1, -1, -1), e1 = ng, e2 = ng, bk = ng;
    float rad = 0.f;
    int msk = 0;
    if (inrange) {
        p4 = A.pos[ia];
        q = A.quat[ia];
        mom = A.dpos_mom[ia];
        qmom = A.dquat_mom[ia];
        ng = A.neighs[lid];
        e1 = A.excl1[lid];
        e2 = A.excl2[lid];
        bk = A.bk_slots[lid];
        rad = A.radius[lid];
        msk = A.fixmask ? A.fixmask[lid] : 0;
    }
    const float invM = p4.w;
    const bool active = inrange && invM > 1e-12f;
    const bool isnode = inrange && lid < nnode;
    const int ngs[4] = {ng.x, ng.y, ng.z, ng.w};
    const float mi = active ? 1.0f / invM : 0.f;
    const float invI = A.rot_mass_scale / (0.4f * mi + 1e-12f);
    const float dt2 = A.dt * A.dt + 1e-16f;
    const int bks[4] = {bk.x, bk.y, bk.z, bk.w};

    // relax mode: replicas still running
    bool running = sys < A.nsys;
    int steps_done = 0;
    if (A.relax && running) {
        steps_done = A.steps[sys];
        running = A.status[sys] == SMR_RAFF_RUNNING;
    }
    float4 dcoll = make_float4(0.f, 0.f, 0.f, 0.f), dnode = dcoll, drot = dcoll;
    // a replica that converged in an earlier launch keeps its buffers and metric
    float metric = (A.relax && sys < A.nsys) ? A.metric[sys] : 0.f;
    bool ran = false;
    float4 *slot = l_slot + r * SEG * 4;

    // SPREAD: the replica's ports (slots k of node i with a neighbour j < n
    // and a positive stiffness, node-major) go to the lanes of its segment,
    // one each (a table in LDS, made once per launch); a node adds its ports'
    // contributions in slot order, as the loop over its slots does.  Port
    // slots without a port are never written: they stay zero.  With more
    // ports than lanes the node loop runs instead.
    bool spread = false;
    int pfirst = 0, pcnt = 0;
    const int *ptab = l_port + r * SEG;
    const float4 *parm = l_parm + r * SEG;
    if constexpr (SPREAD) {
        // each node lane lists its own ports; an exclusive prefix over the
        // segment places them (node-major, slot order)
        float kk[4] = {0.f, 0.f, 0.f, 0.f};
        int cnt = 0;
        if (lid < nnode && lid < n) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                kk[k] = A.kflat[lid * 4 + k];
                cnt += (ngs[k] >= 0 && ngs[k] < n && kk[k] > 0.f) ? 1 : 0;
            }
        }
        int first = cnt;
#pragma unroll
        for (int o = 1; o < SEG; o <<= 1) {
            const int y = __shfl_up(first, o, SEG);
            if (lid >= o) first += y;
        }
        const int total = __shfl(first, SEG - 1, SEG);
        first -= cnt;
        if (lid < nnode && lid < n) {
            int c = first;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (ngs[k] >= 0 && ngs[k] < n && kk[k] > 0.f) {
                    if (c < SEG) {
                        const float4 a = A.port_local[lid * 4 + k];
                        l_parm[r * SEG + c] = make_float4(a.x, a.y, a.z, kk[k]);
                        l_port[r * SEG + c] = (lid * 4 + k) | (ngs[k] << 16);
                    }
                    c++;
                }
            l_pfirst[r * SEG + lid] = first | (cnt << 16);
        }
        if (lid == 0) l_nport[r] = total;
        for (int e = lid; e < SEG * 4; e += SEG) slot[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        spread = l_nport[r] <= SEG;
        if (spread && isnode) {
            const int pf = l_pfirst[r * SEG + lid];
            pfirst = pf & 0xffff;
            pcnt = pf >> 16;
        }
    }

    for (int it = 0; it < A.nsteps; it++) {
        // does any replica of this workgroup still run (relax mode)?
        if (A.relax) {
            if (lid == 0) l_run[r] = running ? 1 : 0;
            __syncthreads();
            int any = 0;
#pragma unroll
            for (int q2 = 0; q2 < R; q2++) any |= l_run[q2];
            __syncthreads();
            if (!any) break;
        }
        l_pos[tid] = p4;
        l_rad[tid] = (active && rad > 0.f) ? rad : 0.f;
        if constexpr (SPREAD) l_quat[tid] = q;
        __syncthreads();
        const bool work = running && active;
        // ---- collision (compute_collision_multi) ----
        float3 sc = make_float3(0.f, 0.f, 0.f);
        if (A.collide && work && rad > 0.f) {
            const float3 p = v3(p4);
            const float4 *rp = l_pos + r * SEG;
            const float *rr = l_rad + r * SEG;
            for (int j = 0; j < n; j++) {
                if (j == lid || excluded(j, e1, e2)) continue;
                const float4 pj4 = rp[j];
                const float rj = rr[j];
                if (pj4.w <= 1e-12f || rj <= 0.f) continue;
                const float3 d = vsub(p, v3(pj4));
                const float d2 = vdot(d, d);
                const float rs = rad + rj;
                if (d2 < rs * rs && d2 > 1e-16f) {
                    const float dist = sqrtf(d2);
                    const float3 nn = vmul(d, 1.0f / dist);
                    const float wt = invM + pj4.w + 1e-12f;
                    const float dl = 0.5f * ((rs - dist) / wt);
                    sc = vadd(sc, vmul(nn, dl * invM));
                }
            }
        }
        // ---- ports (compute_ports_current_multi) ----
        float3 sdp = make_float3(0.f, 0.f, 0.f), sdt = make_float3(0.f, 0.f, 0.f);
        if (SPREAD && spread) {
            // one port per lane: the node's state from LDS (its invI and
            // activity recomputed with the node lane's operations), the
            // contributions into the port's entry, the recoil into its slot
            const float4 *rp = l_pos + r * SEG;
            const int pp = lid;
            if (pp < l_nport[r]) {
                const int code = ptab[pp];
                const int i = (code & 0xffff) >> 2, k = code & 3, j = code >> 16;
                const float4 arm4 = parm[pp];
                const float4 xi4 = rp[i];
                const float invMi = xi4.w;
                const bool wi = running && invMi > 1e-12f;
                float4 rec = make_float4(0.f, 0.f, 0.f, 0.f), cdp = rec, cdt = rec;
                if (wi) {
                    const float mii = 1.0f / invMi;
                    const float invIi = A.rot_mass_scale / (0.4f * mii + 1e-12f);
                    const float4 pj4 = rp[j];
                    if (pj4.w > 1e-12f) {
                        const float3 ra = qrot(l_quat[r * SEG + i], v3(arm4));
                        const float3 diff = vsub(v3(pj4), vadd(v3(xi4), ra));
                        const float d2 = vdot(diff, diff);
                        if (d2 >= 1e-16f) {
                            const float dist = sqrtf(d2);
                            const float3 nn = vmul(diff, 1.0f / dist);
                            const float3 rxn = vcross(ra, nn);
                            const float wang = vdot(rxn, rxn) * invIi;
                            const float alpha = 1.0f / (arm4.w * dt2);
                            const bool jnode = j < nnode;
                            float wtot = invMi + pj4.w + wang + alpha + 1e-12f;
                            if (jnode) wtot += wang;
                            float imp = dist / wtot;
                            if (jnode) imp *= 0.5f;
                            const float3 P = vmul(nn, imp);
                            const float3 a = vmul(P, invMi), b = vmul(vcross(ra, P), invIi);
                            cdp = make_float4(a.x, a.y, a.z, 0.f);
                            cdt = make_float4(b.x, b.y, b.z, 0.f);
                            const float3 rj = vmul(P, -pj4.w);
                            rec = make_float4(rj.x, rj.y, rj.z, 0.f);
                        }
                    }
                }
                if (running) {
                    l_pdp[r * SEG + pp] = cdp;
                    l_pdt[r * SEG + pp] = cdt;
                    slot[i * 4 + k] = rec;
                }
            }
            __syncthreads();
            for (int t = 0; t < pcnt; t++) {
                sdp = vadd(sdp, v3(l_pdp[r * SEG + pfirst + t]));
                sdt = vadd(sdt, v3(l_pdt[r * SEG + pfirst + t]));
            }
        } else if (isnode) {
            const float3 xi = v3(p4);
            const float4 *rp = l_pos + r * SEG;
            // this node's port arms and stiffnesses (shared topology, read
            // through the caches)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
                const int j = ngs[k];
                const float kstk = A.kflat[lid * 4 + k];
                if (work && j >= 0 && j < n && kstk > 0.f) {
                    const float4 pj4 = rp[j];
                    if (pj4.w > 1e-12f) {
                        const float3 ra = qrot(q, v3(A.port_local[lid * 4 + k]));
                        const float3 diff = vsub(v3(pj4), vadd(xi, ra));
                        const float d2 = vdot(diff, diff);
                        if (d2 >= 1e-16f) {
                            const float dist = sqrtf(d2);
                            const float3 nn = vmul(diff, 1.0f / dist);
                            const float3 rxn = vcross(ra, nn);
                            const float wang = vdot(rxn, rxn) * invI;
                            const float alpha = 1.0f / (kstk * dt2);
                            const bool jnode = j < nnode;
                            float wtot = invM + pj4.w + wang + alpha + 1e-12f;
                            if (jnode) wtot += wang;
                            float imp = dist / wtot;
                            if (jnode) imp *= 0.5f;
                            const float3 P = vmul(nn, imp);
                            sdp = vadd(sdp, vmul(P, invM));
                            sdt = vadd(sdt, vmul(vcross(ra, P), invI));
                            const float3 rj = vmul(P, -pj4.w);
                            rec = make_float4(rj.x, rj.y, rj.z, 0.f);
                        }
                    }
                }
                if (running) slot[lid * 4 + k] = rec;
            }
        }
        __syncthreads();
        // ---- corrections (apply_corrections_multi) ----
        if (work) {
            float3 dxp = isnode ? sdp : make_float3(0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (bks[k] >= 0) dxp = vadd(dxp, v3(slot[bks[k]]));
            float3 dx = vadd(sc, dxp);
            if (msk & 1) dx.x = 0.f;
            if (msk & 2) dx.y = 0.f;
            if (msk & 4) dx.z = 0.f;
            float3 mv = vmul(dx, A.relaxation);
            if (A.beta != 0.f) mv = vadd(mv, vmul(v3(mom), A.beta));
            if (msk & 8) {
                if (fabsf(p4.z) > 1e-6f) mv.z -= p4.z;
                else mv.z = 0.f;
            }
            p4.x += mv.x;
            p4.y += mv.y;
            p4.z += mv.z;
            mom = make_float4(mv.x, mv.y, mv.z, 0.f);
            if (isnode && !(msk & 7)) {
                const float3 dth = vmul(sdt, A.relaxation);
                const float ang = sqrtf(vdot(dth, dth));
                const float4 qold = q;
                float4 qn = qold;
                if (ang > 1e-8f) qn = qnormalize(qmul(qaxis(vmul(dth, 1.0f / ang), ang), qold));
                if (A.beta != 0.f)
                    qn = make_float4(qn.x + qmom.x * A.beta, qn.y + qmom.y * A.beta, qn.z + qmom.z * A.beta,
                                     qn.w + qmom.w * A.beta);
                qn = qnormalize(qn);
                if (qn.x * qold.x + qn.y * qold.y + qn.z * qold.z + qn.w * qold.w < 0.f)
                    qn = make_float4(-qn.x, -qn.y, -qn.z, -qn.w);
                q = qn;
                qmom = make_float4(qn.x - qold.x, qn.y - qold.y, qn.z - qold.z, qn.w - qold.w);
            }
        }
        // this step's correction buffers (what the reference leaves in
        // dpos_coll / dpos_node / drot_node) and the replica's max |component|
        if (running) {
            ran = true;
            dcoll = make_float4(sc.x, sc.y, sc.z, 0.f);
            dnode = make_float4(sdp.x, sdp.y, sdp.z, 0.f);
            drot = make_float4(sdt.x, sdt.y, sdt.z, 0.f);
        }
        float m = 0.f;
        if (running && inrange) {
            m = fmaxf(fabsf(sc.x), fmaxf(fabsf(sc.y), fabsf(sc.z)));
            if (isnode) m = fmaxf(m, fmaxf(fabsf(sdp.x), fmaxf(fabsf(sdp.y), fabsf(sdp.z))));
        }
        if constexpr (SEG <= 64) {
            m = seg_max<SEG>(m);
        } else {
            m = seg_max<64>(m);
            if ((tid & 63) == 0) l_red[tid >> 6] = m;
            __syncthreads();
            m = l_red[0];
#pragma unroll
            for (int w = 1; w < NW; w++) m = fmaxf(m, l_red[w]);
        }
        if (running) {
            metric = m;
            steps_done++;
            if (A.relax) {
                if (m < A.tol) {
                    running = false;
                    if (lid == 0) A.status[sys] = SMR_RAFF_CONVERGED;
                } else if (steps_done >= A.max_steps) {
                    running = false;
                    if (lid == 0) A.status[sys] = SMR_RAFF_MAXSTEPS;
                }
            }
        }
        __syncthreads();
    }
    if (inrange) {
        A.pos[ia] = p4;
        A.quat[ia] = q;
        A.dpos_mom[ia] = mom;
        A.dquat_mom[ia] = qmom;
    }
    if (inrange && ran) {
        A.dpos_coll[ia] = dcoll;
        if (lid < nnode) {
            const size_t in = (size_t)sys * nnode + lid;
            A.dpos_node[in] = dnode;
            A.drot_node[in] = drot;
#pragma unroll
            for (int k = 0; k < 4; k++) A.dpos_neigh[in * 4 + k] = slot[lid * 4 + k];
        }
    }
    if (sys < A.nsys && lid == 0) {
        A.metric[sys] = metric;
        if (A.relax) A.steps[sys] = steps_done;
    }
}

hipError_t launch_raff(const RaffArgs &A, int seg, hipStream_t st) {
    const int blk = seg < 64 ? 64 : seg, rep = blk / seg;
    const int grid = (A.nsys + rep - 1) / rep;
    switch (seg) {
        case 16: hipLaunchKernelGGL(k_raff<16>, dim3(grid), dim3(blk), 0, st, A); break;
        case 32: hipLaunchKernelGGL(k_raff<32>, dim3(grid), dim3(blk), 0, st, A); break;
        case 64: hipLaunchKernelGGL(k_raff<64>, dim3(grid), dim3(blk), 0, st, A); break;
        case 128: hipLaunchKernelGGL(k_raff<128>, dim3(grid), dim3(blk), 0, st, A); break;
        case 256: hipLaunchKernelGGL(k_raff<256>, dim3(grid), dim3(blk), 0, st, A); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace smr
