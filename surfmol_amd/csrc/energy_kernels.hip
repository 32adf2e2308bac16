// energy_kernels.hip -- FP64 energy evaluation of systems at their current
// (FP32) positions.
//
// The relaxation kernels integrate with FP32 forces and never need the
// energy.  The energies a caller receives (smr_batch_eval_forces,
// smr_relax_*) are evaluated here once, after the forces, in FP64 from the
// FP64 force-field parameters: near a minimum the terms are small
// differences of large numbers (k (l - l0)^2, K (c0 + c2 cos 2t + c3 cos 3t),
// steep LJ walls), which FP32 cannot resolve to 1e-6 eV/atom.
//
// One workgroup per listed system.  Term formulas, and the bookkeeping of
// the reference (bond and pair energies counted at both ends, the angle
// energy without K c1 cos t), follow:
//   bond       molff/src/uff.rs:440-468
//   angle      molff/src/uff.rs:471-507
//   dihedral   molff/src/uff.rs:510-545
//   inversion  molff/src/uff.rs:548-579
//   pair       opencl/UFF.cl:57-95 (getNonBond convention)
//   surface    opencl/gridff_eval.cl:48-69,234-286
#include "smr_device.h"
#include "smr_kernels.h"

namespace smr {

namespace {

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 mk(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 add(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 mul(d3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ d3 ld(const float4 *p, int i) {
    const float4 v = p[i];
    return mk((double)v.x, (double)v.y, (double)v.z);
}

// unit vector of device hneigh slot g = k*N + atom (atom -> its k-th neighbour)
__device__ __forceinline__ d3 slot_dir(const float4 *pos, const int4 *ng, int g, int N) {
    const int k = g / N;
    const int a = g - k * N;
    const int4 n = ng[a];
    const int b = k == 0 ? n.x : k == 1 ? n.y : k == 2 ? n.z : n.w;
    d3 d = sub(ld(pos, b), ld(pos, a));
    return mul(d, 1.0 / sqrt(dot(d, d)));
}

__device__ __forceinline__ void basis64(double u, double *b) {
    const double inv6 = 1.0 / 6.0;
    double u2 = u * u, t = 1.0 - u;
    b[0] = inv6 * t * t * t;
    b[1] = inv6 * (3.0 * u2 * (u - 2.0) + 4.0);
    b[2] = inv6 * (3.0 * u * (1.0 + u - u2) + 1.0);
    b[3] = inv6 * u2 * u;
}

}  // namespace

__global__ __launch_bounds__(256) void k_energy64(DevTables T, const float4 *__restrict__ pos,
                                                  SysState *__restrict__ state, const int *__restrict__ list) {
    const int s = list[blockIdx.x];
    const SysDesc sd = T.sys[s];
    const TopoDesc td = T.topo[sd.topo];
    const int N = td.natoms;
    const float4 *P = pos + sd.atom_base;
    const int4 *NG = T.ng + td.atom_off;
    const int tid = threadIdx.x;
    double e[6] = {0, 0, 0, 0, 0, 0};
    // bonds, counted at both ends (uff.rs:462)
    for (int i = tid; i < N; i += blockDim.x) {
        const int4 n = NG[i];
        const int ns[4] = {n.x, n.y, n.z, n.w};
        const double4 k = T.bk64[td.atom_off + i], l0 = T.bl64[td.atom_off + i];
        const double ks[4] = {k.x, k.y, k.z, k.w}, ls[4] = {l0.x, l0.y, l0.z, l0.w};
        const d3 pi = ld(P, i);
        for (int q = 0; q < 4; q++) {
            if (ns[q] < 0) continue;
            d3 d = sub(ld(P, ns[q]), pi);
            const double dl = sqrt(dot(d, d)) - ls[q];
            e[0] += ks[q] * dl * dl;
        }
    }
    // angles: E = K (c0 + c2 cos 2t + c3 cos 3t)
    for (int a = tid; a < td.nangles; a += blockDim.x) {
        const unsigned g = T.ang_ng[td.ang_off + a];
        const d3 qij = slot_dir(P, NG, (int)(g & 0xffffu), N), qkj = slot_dir(P, NG, (int)(g >> 16), N);
        const double4 p = T.ang64[td.ang_off + a];
        const d3 h = add(qij, qkj);
        const double c = 0.5 * (dot(h, h) - 2.0);
        const double sn = sqrt(fmax(0.0, 1.0 - c * c + 1e-14));
        const double cx = c * c - sn * sn, cy = 2.0 * c * sn;
        const double c3x = cx * c - cy * sn;
        e[1] += p.x * (p.y + p.z * cx + p.w * c3x);
    }
    // dihedrals: E = V (1 + d cos n phi)
    for (int q = tid; q < td.ndih; q += blockDim.x) {
        const uint2 g = T.dih_ng[td.dih_off + q];
        const d3 h12 = slot_dir(P, NG, (int)(g.x & 0xffffu), N), h32 = slot_dir(P, NG, (int)(g.x >> 16), N);
        const d3 h43 = slot_dir(P, NG, (int)(g.y & 0xffffu), N);
        const double4 p = T.dih64[td.dih_off + q];
        const d3 n123 = cross(h12, h32), n234 = cross(h43, h32);
        const double n1 = dot(n123, n123), n2 = dot(n234, n234);
        if (n1 < 1e-30 || n2 < 1e-30) continue;
        const double inv_n12 = sqrt(1.0 / n1 * (1.0 / n2));
        const double csx = dot(n123, n234) * inv_n12, csy = -dot(n123, h43) * inv_n12;
        double cnx = csx, cny = csy;
        const int n = (int)p.z;
        for (int i = 1; i < n; i++) {
            const double tx = cnx * csx - cny * csy;
            cny = cnx * csy + cny * csx;
            cnx = tx;
        }
        e[2] += p.x * (1.0 + p.y * cnx);
    }
    // inversions: E = K (C0 + C1 c + C2 cos 2w)
    for (int q = tid; q < td.ninv; q += blockDim.x) {
        const uint2 g = T.inv_ng[td.inv_off + q];
        const d3 h21 = slot_dir(P, NG, (int)(g.x & 0xffffu), N), h31 = slot_dir(P, NG, (int)(g.x >> 16), N);
        const d3 h41 = slot_dir(P, NG, (int)g.y, N);
        const double4 p = T.inv64[td.inv_off + q];
        d3 n123 = cross(h21, h31);
        const double nl = sqrt(dot(n123, n123));
        n123 = nl > 1e-8 ? mul(n123, 1.0 / nl) : mk(0.0, 0.0, 0.0);
        const double sn = -dot(n123, h41);
        const double c = sqrt(fmax(0.0, 1.0 - sn * sn + 1e-14));
        const double c2x = c * c - sn * sn;
        e[3] += p.x * (p.y + p.z * c + p.w * c2x);
    }
    // pairs: every unordered non-bonded pair once, counted at both ends.
    // Pair p of the block is atom i with i + d (mod N) on sweep d = p / N + 1
    // (d = 1..N/2; the last sweep of an even N takes i < N/2 only), so the
    // block's threads share the N(N-1)/2 pairs evenly.
    if (T.use_pairs) {
        const int nd = N >> 1;
        double acc = 0.0;
        for (int p = tid; p < N * nd; p += blockDim.x) {
            const int d = p / N + 1, i = p - (d - 1) * N;
            if ((N & 1) == 0 && d == nd && i >= nd) continue;
            int j = i + d;
            if (j >= N) j -= N;
            const int4 n = NG[i];
            if (j == n.x || j == n.y || j == n.z || j == n.w) continue;
            const double4 ri = T.req64[td.atom_off + i], rj = T.req64[td.atom_off + j];
            const d3 dp = sub(ld(P, j), ld(P, i));
            const double r2 = dot(dp, dp);
            const double R0 = ri.x + rj.x, E0 = ri.y * rj.y, kQ = ri.z * rj.z;
            const double u2 = R0 * R0 / r2;
            const double u6 = u2 * u2 * u2;
            acc += (u6 - 2.0) * u6 * E0 + kQ / sqrt(r2 + T.r2damp64);
        }
        e[4] += 2.0 * acc;
    }
    // surface: E = sum_stencil Bx By Bz (PLQH . C); the block's threads
    // share the (atom, stencil column) items, 16 per atom
    if (T.grid) {
        for (int it = tid; it < 16 * N; it += blockDim.x) {
            const int i = it >> 4, a = (it >> 2) & 3, b = it & 3;
            const d3 p = ld(P, i);
            const double ux = p.x * T.ginv64_x, uy = p.y * T.ginv64_y, uz = p.z * T.ginv64_z;
            const double fx = floor(ux), fy = floor(uy);
            const int iz = (int)uz;
            if (iz < 1 || iz >= T.gnz - 2) continue;
            double bx[4], by[4], bz[4];
            basis64(ux - fx, bx);
            basis64(uy - fy, by);
            basis64(uz - (double)iz, bz);
            int ix = sd.iux + (int)fx - 1, iy = sd.iuy + (int)fy - 1;
            ix %= T.gnx; if (ix < 0) ix += T.gnx;
            iy %= T.gny; if (iy < 0) iy += T.gny;
            int xi = ix + a; if (xi >= T.gnx) xi -= T.gnx;
            int yi = iy + b; if (yi >= T.gny) yi -= T.gny;
            const float4 *col = T.grid + ((size_t)(xi * T.gny + yi) * T.gnz + (iz - 1));
            const double4 w = T.plq64[td.atom_off + i];
            double ez = 0.0;
            for (int c = 0; c < 4; c++) {
                const float4 C = col[c];
                ez += bz[c] * (w.x * (double)C.x + w.y * (double)C.y + w.z * (double)C.z);
            }
            e[5] += bx[a] * by[b] * ez;
        }
    }
    // block reduction (wave shuffles, then waves in order)
    __shared__ double red[4][6];
    const int lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const double r = wave_sum(e[q]);
        if (lane == 0) red[wid][q] = r;
    }
    __syncthreads();
    if (tid == 0) {
        for (int q = 0; q < 6; q++) {
            double acc = 0.0;
            for (int w = 0; w < nw; w++) acc += red[w][q];
            state[s].E[q] = acc;
        }
    }
}

hipError_t launch_energy64(const DevTables &T, const float4 *pos, SysState *state, const int *list, int n,
                           hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_energy64, dim3(n), dim3(256), 0, st, T, pos, state, list);
    return hipGetLastError();
}

}  // namespace smr
