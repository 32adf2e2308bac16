// smr_device.h -- device-side data layout and force-field math shared by the
// fused relax kernel and the per-term kernels of the plain method.
//
// Formulas follow the reference (paths relative to the reference root):
//   bond       molff/src/uff.rs:440-468      opencl/UFF.cl:141-322
//   angle      molff/src/uff.rs:471-507      opencl/UFF.cl:326-501
//   dihedral   molff/src/uff.rs:510-545      opencl/UFF.cl:505-670
//   inversion  molff/src/uff.rs:548-579      opencl/UFF.cl:674-795
//   pair       opencl/UFF.cl:57-95 (mixREQ_arithmetic + getLJQH)
//   surface    opencl/gridff_eval.cl:48-69,234-286; opencl/UFF.cl:1673-1711
// with the guards of the OpenCL versions where the FP64 CPU code has none
// (clamped cosines, degenerate-normal checks).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SMR_COULOMB_CONST 14.3996448915f

// ---------------------------------------------------------------------------
// packed tables (built on the host by capi.cpp)
// ---------------------------------------------------------------------------
struct TopoDesc {
    int natoms, nangles, ndih, ninv;
    int nf;          // 4*ndih + 4*ninv + 3*nangles force pieces
    int atom_off;    // into the per-atom pools
    int ang_off, dih_off, inv_off;
    int a2f_off;     // into the piece-index pool
};

struct SysDesc {
    int topo;        // index into TopoDesc[]
    int atom_base;   // into the state arrays (pos/vel/force)
    int iux, iuy;    // grid-node offset of the system origin, wrapped into [0,nx),[0,ny)
    int gx, gy, gz;  // unwrapped origin offset in origin-grid steps (org = p0 + g*step)
    int pad;
};

// FIRE / MD state kept per system across launches
struct SysState {
    double dt, alpha;
    int npos, steps, status, pad;
    double fmax;
    double E[6];     // bond, angle, dihedral, inversion, pair, surface
};

struct FireConst {
    double dt0, dt_max, alpha0, f_inc, f_dec, f_alpha, flim, fconv;
    int n_min, max_steps;
    // damped MD (mode KMODE_MD)
    double md_dt, md_flim, md_cdamp;
};

struct DevTables {
    const TopoDesc *topo;
    const SysDesc *sys;
    // per-atom pools (index TopoDesc.atom_off + i)
    const int4 *ng;        // neighbour atoms (local), -1 padded
    const float4 *bk;      // bond k per neighbour slot
    const float4 *bl;      // bond l0 per neighbour slot
    const float4 *req;     // (R, sqrtE, Q, 0)
    const float4 *plq;     // (e^{2aR} sqrtE, e^{aR} sqrtE, Q, 0)
    const int2 *a2f;       // (start, count) into the piece-index pool (relative to a2f_off)
    // per-interaction pools; hneigh slots (4*atom+slot, local) packed 16-bit
    const unsigned *ang_ng;      // slot_ji | slot_jk << 16
    const float4 *ang_par;       // K*(c0, c1, c2, c3)
    const uint2 *dih_ng;         // (s12 | s32 << 16, s43 | n << 16), n = multiplicity
    const float4 *dih_par;       // (V, V*d, V*d*n, 0)
    const uint2 *inv_ng;         // (s21 | s31 << 16, s41)
    const float4 *inv_par;       // K*(C0, C1, C2), 0
    const unsigned short *a2f_idx;  // piece indices (local to the topology's fint layout)
    // where each force piece of a term goes in the atom-major piece order
    // (the a2f lists inverted; 16-bit local positions, 4 per uint2, piece k
    // in bits 16*(k&1) of component k>>1)
    const uint2 *ang_dst, *dih_dst, *inv_dst;
    // GridFF
    const float4 *grid;    // [nx*ny*nz] z-fastest (P, L, Q, 0); NULL = no surface
    int gnx, gny, gnz;
    float ginv_x, ginv_y, ginv_z;
    // origin grid of the per-system local frames (the GridFF nodes, or 1 A)
    float ostep_x, ostep_y, ostep_z;
    double ostep64_x, ostep64_y, ostep64_z;
    // options
    float r2damp;
    int use_pairs;
    // FP64 parameters of the energy evaluation (k_energy64): per atom bond k
    // and l0 per neighbour slot, (R, sqrtE, sqrt(kC) Q, 0), PLQH; per term
    // angle (K, c0, c2, c3), dihedral (V, d, n, 0), inversion (K, C0, C1, C2)
    const double4 *bk64, *bl64, *req64, *plq64;
    const double4 *ang64, *dih64, *inv64;
    double ginv64_x, ginv64_y, ginv64_z;
    double r2damp64;
};

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float3 f3(float4 a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// a*u + b*w (Vec3d::set_lincomb)
__device__ __forceinline__ float3 lincomb(float a, float3 u, float b, float3 w) {
    return make_float3(a * u.x + b * w.x, a * u.y + b * w.y, a * u.z + b * w.z);
}

// 1-ulp hardware reciprocal and reciprocal square root (v_rcp_f32,
// v_rsq_f32) instead of the IEEE-exact division / sqrt sequences.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Compensated (Kahan) accumulator for one force vector.
struct KahanF3 {
    float3 s, c;
    __device__ __forceinline__ void init(float3 v) { s = v; c = make_float3(0.f, 0.f, 0.f); }
    __device__ __forceinline__ void add1(float &sv, float &cv, float x) {
        float y = x - cv;
        float t = sv + y;
        cv = (t - sv) - y;
        sv = t;
    }
    __device__ __forceinline__ void add(float3 x) { add1(s.x, c.x, x.x); add1(s.y, c.y, x.y); add1(s.z, c.z, x.z); }
    __device__ __forceinline__ float3 value() const { return make_float3(s.x - c.x, s.y - c.y, s.z - c.z); }
};

// ---------------------------------------------------------------------------
// force-field terms
// ---------------------------------------------------------------------------

// Bond between atom i (at pi) and neighbour at pj: writes hneigh (h, 1/l),
// adds 2k(l-l0) h to f, returns k (l-l0)^2.   uff.rs:440-468 / UFF.cl:223-256
// The stretch l - l0 is formed as (|d|^2 - l0^2) / (l + l0) with |d|^2 and
// the subtraction in FP64 from the exact difference of the FP32 positions:
// l - l0 taken from an FP32 length carries an error of ~1e-7 l (not of
// l - l0), i.e. ~1e-5 eV/A of force on a stiff bond whatever the strain.
__device__ __forceinline__ float bond_term(float3 pi, float3 pj, float k, float l0, float4 &hn, float3 &f) {
    float3 dp = pj - pi;
    const double dx = (double)pj.x - (double)pi.x, dy = (double)pj.y - (double)pi.y,
                 dz = (double)pj.z - (double)pi.z;
    const double l2d = fma(dx, dx, fma(dy, dy, dz * dz));
    const double l0d = (double)l0;
    float l2 = fmaxf((float)l2d, 1e-16f);
    float inv_l = frsq(l2);
    float l = l2 * inv_l;
    float3 h = dp * inv_l;
    hn = make_float4(h.x, h.y, h.z, inv_l);
    float dl = (float)fma(-l0d, l0d, l2d) * frcp(fmaxf(l + l0, 1e-8f));
    f = f + h * (2.0f * k * dl);
    return k * dl * dl;
}

// Angle (Fourier, complex powers): uff.rs:471-507 / UFF.cl:398-447.
// kc = K*(c0, c1, c2, c3).  Energy omits K*c1*cos(theta) as the reference does.
__device__ __forceinline__ float angle_term(float4 qij, float4 qkj, float4 kc, float3 &fpi, float3 &fpj, float3 &fpk) {
    float3 a = f3(qij), b = f3(qkj);
    float3 h = a + b;
    float c = 0.5f * (dot3(h, h) - 2.0f);
    c = fminf(fmaxf(c, -1.0f), 1.0f);
    float s2 = fmaxf(0.0f, 1.0f - c * c) + 1e-14f;
    float inv_s = frsq(s2);
    float s = s2 * inv_s;
    float e = kc.x, ff = kc.y;
    float cx = c * c - s * s, cy = 2.0f * c * s;          // (cos2t, sin2t)
    e += kc.z * cx;
    ff += kc.z * cy * inv_s * 2.0f;
    float c3x = cx * c - cy * s, c3y = cx * s + cy * c;   // (cos3t, sin3t)
    e += kc.w * c3x;
    ff += kc.w * c3y * inv_s * 3.0f;
    float fi = ff * qij.w, fk = ff * qkj.w;
    float fic = fi * c, fkc = fk * c;
    fpi = lincomb(fic, a, -fi, b);
    fpk = lincomb(-fk, a, fkc, b);
    fpj = lincomb(fk - fic, a, fi - fkc, b);
    return e;
}

// Dihedral: uff.rs:510-545 / UFF.cl:563-603.  par = (V, V*d, V*d*n, 0), n = multiplicity.
__device__ __forceinline__ float dihedral_term(float4 q12, float4 q32, float4 q43, float4 par, int n, float3 &fp1,
                                               float3 &fp2, float3 &fp3, float3 &fp4) {
    float3 h12 = f3(q12), h32 = f3(q32), h43 = f3(q43);
    float3 n123 = cross3(h12, h32);
    float3 n234 = cross3(h43, h32);
    float n1 = dot3(n123, n123), n2 = dot3(n234, n234);
    if (n1 < 1e-30f || n2 < 1e-30f) {
        fp1 = fp2 = fp3 = fp4 = make_float3(0.f, 0.f, 0.f);
        return 0.0f;
    }
    float il2_123 = frcp(n1), il2_234 = frcp(n2);
    float inv_n12 = frsq(n1) * frsq(n2);
    float csx = dot3(n123, n234) * inv_n12;
    float csy = -dot3(n123, h43) * inv_n12;
    float cnx = csx, cny = csy;
    for (int i = 1; i < n; i++) {
        float tx = cnx * csx - cny * csy;
        float ty = cnx * csy + cny * csx;
        cnx = tx; cny = ty;
    }
    float e = par.x + par.y * cnx;
    float f = -par.z * cny;
    fp1 = n123 * (-f * il2_123 * q12.w);
    fp4 = n234 * (f * il2_234 * q43.w);
    float c123 = dot3(h32, h12) * (q32.w * frcp(q12.w));
    float c432 = dot3(h32, h43) * (q32.w * frcp(q43.w));
    fp3 = lincomb(-c123, fp1, -c432 - 1.0f, fp4);
    fp2 = lincomb(c123 - 1.0f, fp1, c432, fp4);
    return e;
}

// Inversion: uff.rs:548-579 / UFF.cl:706-750.  par = K*(C0, C1, C2), 0.
__device__ __forceinline__ float inversion_term(float4 q21, float4 q31, float4 q41, float4 par, float3 &fp1, float3 &fp2,
                                                float3 &fp3, float3 &fp4) {
    float3 h21 = f3(q21), h31 = f3(q31), h41 = f3(q41);
    float3 n123 = cross3(h21, h31);
    float nn = dot3(n123, n123);
    float il123 = 0.0f;
    if (nn > 1e-16f) {
        il123 = rsqrtf(nn);
        n123 = n123 * il123;
    } else {
        n123 = make_float3(0.f, 0.f, 0.f);
    }
    float s = -dot3(n123, h41);
    s = fminf(fmaxf(s, -1.0f), 1.0f);
    float cc = fmaxf(0.0f, 1.0f - s * s) + 1e-14f;
    float inv_c = frsq(cc);
    float c = cc * inv_c;
    float c2x = c * c - s * s, c2y = 2.0f * c * s;
    float e = par.x + par.y * c + par.z * c2x;
    float f = (c > 1e-7f) ? -(par.y * s + 2.0f * par.z * c2y) * inv_c : 0.0f;
    float fq41 = f * q41.w;
    float fi123 = f * il123;
    float3 tq = lincomb(s * fi123, n123, fi123, h41);
    fp4 = lincomb(fq41, n123, s * fq41, h41);
    fp2 = cross3(h31, tq) * q21.w;
    fp3 = cross3(tq, h21) * q31.w;
    fp1 = (fp2 + fp3 + fp4) * -1.0f;
    return e;
}

// getLJQH (UFF.cl:82-95) after mixREQ_arithmetic (UFF.cl:57-65):
// R0 = Ri+Rj, E0 = sqrtEi*sqrtEj, kC*Q = qi*qj where the staged REQ rows
// hold q = sqrt(kC)*Q.  dp = pj - pi; returns E and the force on i in f.
__device__ __forceinline__ float ljq_pair(float3 dp, float4 ri, float4 rj, float r2damp, float3 &f) {
    float R0 = ri.x + rj.x;
    float E0 = ri.y * rj.y;
    float kQ = ri.z * rj.z;
    float r2 = dot3(dp, dp);
    float ir_ = frsq(r2 + r2damp);
    float ir2_ = ir_ * ir_;
    float Ec = kQ * ir_;
    float ir2 = frcp(r2);
    float u2 = R0 * R0 * ir2;
    float u6 = u2 * u2 * u2;
    float vdW = u6 * E0;
    float E = (u6 - 2.0f) * vdW + Ec;
    float fr = -12.0f * (u6 - 1.0f) * vdW * ir2 - Ec * ir2_;
    f = dp * fr;
    return E;
}

// The same pair force as a radial factor: the force on i is dp * fr (no energy).
__device__ __forceinline__ float ljq_pair_fr(float3 dp, float4 ri, float4 rj, float r2damp) {
    const float R0 = ri.x + rj.x;
    const float E0 = ri.y * rj.y;
    const float kQ = ri.z * rj.z;
    const float r2 = dot3(dp, dp);
    const float ir_ = frsq(r2 + r2damp);
    const float ir2 = frcp(r2);
    const float u2 = R0 * R0 * ir2;
    const float u6 = u2 * u2 * u2;
    const float vdW = u6 * E0;
    return -12.0f * (u6 - 1.0f) * vdW * ir2 - (kQ * ir_) * (ir_ * ir_);
}

// The same radial factor from a REQ row ri whose sqrtE and q carry the
// force's constant factors (ri.y = -12 sqrtEi, ri.z = -qi): one multiply
// fewer per pair.
__device__ __forceinline__ float ljq_pair_fr_folded(float3 dp, float4 ri, float4 rj, float r2damp) {
    const float R0 = ri.x + rj.x;
    const float E0 = ri.y * rj.y;
    const float kQ = ri.z * rj.z;
    const float r2 = dot3(dp, dp);
    const float ir_ = frsq(r2 + r2damp);
    const float ir2 = frcp(r2);
    const float u2 = R0 * R0 * ir2;
    const float u6 = u2 * u2 * u2;
    return (u6 - 1.0f) * (u6 * E0) * ir2 + (kQ * ir_) * (ir_ * ir_);
}

// ---------------------------------------------------------------------------
// GridFF tricubic B-spline sample: fe3d_pbc_comb (gridff_eval.cl:234-286)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 bsp_basis(float u) {
    const float inv6 = 1.0f / 6.0f;
    float u2 = u * u, t = 1.0f - u;
    return make_float4(inv6 * t * t * t, inv6 * (3.0f * u2 * (u - 2.0f) + 4.0f), inv6 * (3.0f * u * (1.0f + u - u2) + 1.0f),
                       inv6 * u2 * u);
}
__device__ __forceinline__ float4 bsp_dbasis(float u) {
    float u2 = u * u, t = 1.0f - u;
    return make_float4(-0.5f * t * t, 0.5f * (3.0f * u2 - 4.0f * u), 0.5f * (-3.0f * u2 + 2.0f * u + 1.0f), 0.5f * u2);
}
__device__ __forceinline__ float dot4(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
// PLQH . C over the three used channels (the fourth channel is zero)
__device__ __forceinline__ float dotplq(float4 p, float4 c) { return p.x * c.x + p.y * c.y + p.z * c.z; }

// Sample at grid coordinate u = (ix0 + ul.x, iy0 + ul.y, ul.z), where (ix0,iy0)
// is the system's integer node offset and ul the local fractional
// coordinate.  Returns (dE/du, E); zero outside 1 <= iz < nz-2.
__device__ __forceinline__ float4 grid_sample(const float4 *__restrict__ G, int nx, int ny, int nz, int ix0, int iy0,
                                              float3 ul, float4 plq) {
    float fx = floorf(ul.x), fy = floorf(ul.y);
    int iz = (int)ul.z;  // truncation, as (int)u.z in the reference
    float tx = ul.x - fx, ty = ul.y - fy, tz = ul.z - (float)iz;
    if (iz < 1 || iz >= nz - 2) return make_float4(0.f, 0.f, 0.f, 0.f);
    int ix = ix0 + (int)fx - 1;
    int iy = iy0 + (int)fy - 1;
    ix %= nx; if (ix < 0) ix += nx;
    iy %= ny; if (iy < 0) iy += ny;
    float4 bz = bsp_basis(tz), dz = bsp_dbasis(tz);
    float4 by = bsp_basis(ty), dy = bsp_dbasis(ty);
    float4 bx = bsp_basis(tx), dx = bsp_dbasis(tx);
    float by_[4] = {by.x, by.y, by.z, by.w}, dy_[4] = {dy.x, dy.y, dy.z, dy.w};
    float bx_[4] = {bx.x, bx.y, bx.z, bx.w}, dx_[4] = {dx.x, dx.y, dx.z, dx.w};
    float ox = 0.f, oy = 0.f, oz = 0.f, oe = 0.f;
#pragma unroll 1
    for (int a = 0; a < 4; a++) {
        int xi = ix + a; if (xi >= nx) xi -= nx;
        float e_y = 0.f, e_dy = 0.f, e_dz = 0.f;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            int yi = iy + b; if (yi >= ny) yi -= ny;
            const float4 *col = G + ((size_t)(xi * ny + yi) * nz + (iz - 1));
            float4 c0 = col[0], c1 = col[1], c2 = col[2], c3 = col[3];
            float4 cs = make_float4(dotplq(plq, c0), dotplq(plq, c1), dotplq(plq, c2), dotplq(plq, c3));
            float fe = dot4(bz, cs);   // fe1Dcomb
            float fd = dot4(dz, cs);
            e_dy += fe * dy_[b];       // fe2d_comb
            e_dz += fd * by_[b];
            e_y += fe * by_[b];
        }
        ox += dx_[a] * e_y;            // fe3d_pbc_comb
        oy += bx_[a] * e_dy;
        oz += bx_[a] * e_dz;
        oe += bx_[a] * e_y;
    }
    return make_float4(ox, oy, oz, oe);
}

// The same sample split in two, for a per-atom stencil cache: locate the
// cell, contract its 4x4x4 stencil with the atom's PLQH once (stored while
// the atom stays in the cell), then interpolate inside the cell.
struct GridCell {
    int ix, iy, iz;     // first stencil node (wrapped x,y); iz < 0: outside
    float tx, ty, tz;
};

// i mod n in [0, n): one add or subtract for |i| within a period of the
// range (the usual case), an integer division only beyond that
__device__ __forceinline__ int wrap_index(int i, int n) {
    if (i >= n) i -= n;
    else if (i < 0) i += n;
    if ((unsigned)i >= (unsigned)n) {
        i %= n;
        if (i < 0) i += n;
    }
    return i;
}

__device__ __forceinline__ GridCell grid_cell(int nx, int ny, int nz, int ix0, int iy0, float3 ul) {
    GridCell g;
    float fx = floorf(ul.x), fy = floorf(ul.y);
    int iz = (int)ul.z;
    g.tx = ul.x - fx;
    g.ty = ul.y - fy;
    g.tz = ul.z - (float)iz;
    if (iz < 1 || iz >= nz - 2) {
        g.ix = g.iy = 0;
        g.iz = -1;
        return g;
    }
    int ix = ix0 + (int)fx - 1;
    int iy = iy0 + (int)fy - 1;
    ix = wrap_index(ix, nx);
    iy = wrap_index(iy, ny);
    g.ix = ix;
    g.iy = iy;
    g.iz = iz - 1;
    return g;
}

// Contract the 4x4x4 stencil of cell g with the atom's PLQH and store it:
// sc[(4a+b)*ld] = (PLQH.C[x+a][y+b][z+c], c = 0..3), 256 B per atom.  The
// stride ld (the system's atom count) makes element k of all atoms of a
// system contiguous, so one load instruction of a wave reads whole lines.
__device__ __forceinline__ void grid_store_stencil(const float4 *__restrict__ G, int nx, int ny, int nz,
                                                   const GridCell &g, float4 plq, float4 *sc, int ld) {
#pragma unroll 1
    for (int a = 0; a < 4; a++) {
        int xi = g.ix + a; if (xi >= nx) xi -= nx;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            int yi = g.iy + b; if (yi >= ny) yi -= ny;
            const float4 *col = G + ((size_t)(xi * ny + yi) * nz + g.iz);
            sc[(4 * a + b) * ld] = make_float4(dotplq(plq, col[0]), dotplq(plq, col[1]), dotplq(plq, col[2]),
                                               dotplq(plq, col[3]));
        }
    }
}

// Contract the 4x4x4 stencil of cell g with the atom's PLQH into registers
// (the same values grid_store_stencil writes).  The scheduling barrier after
// each x plane keeps at most 16 node loads in flight, so the contraction does
// not raise the kernel's register count.
__device__ __forceinline__ void grid_contract_stencil(const float4 *__restrict__ G, int nx, int ny, int nz,
                                                      const GridCell &g, float4 plq, float4 (&rs)[16]) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
        int xi = g.ix + a; if (xi >= nx) xi -= nx;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            int yi = g.iy + b; if (yi >= ny) yi -= ny;
            const float4 *col = G + ((size_t)(xi * ny + yi) * nz + g.iz);
            rs[4 * a + b] = make_float4(dotplq(plq, col[0]), dotplq(plq, col[1]), dotplq(plq, col[2]),
                                        dotplq(plq, col[3]));
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Interpolate inside cell g from a contracted stencil, element (4a+b) given
// by get(4a+b) (the contraction order of grid_sample: z, then y, then x).
// Returns (dE/du, E).
template <class Get>
__device__ __forceinline__ float4 grid_eval_stencil(Get get, const GridCell &g) {
    if (g.iz < 0) return make_float4(0.f, 0.f, 0.f, 0.f);
    float4 bz = bsp_basis(g.tz), dz = bsp_dbasis(g.tz);
    float4 by = bsp_basis(g.ty), dy = bsp_dbasis(g.ty);
    float4 bx = bsp_basis(g.tx), dx = bsp_dbasis(g.tx);
    float by_[4] = {by.x, by.y, by.z, by.w}, dy_[4] = {dy.x, dy.y, dy.z, dy.w};
    float bx_[4] = {bx.x, bx.y, bx.z, bx.w}, dx_[4] = {dx.x, dx.y, dx.z, dx.w};
    float ox = 0.f, oy = 0.f, oz = 0.f, oe = 0.f;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        float e_y = 0.f, e_dy = 0.f, e_dz = 0.f;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const float4 cs = get(4 * a + b);
            float fe = dot4(bz, cs);
            float fd = dot4(dz, cs);
            e_dy += fe * dy_[b];
            e_dz += fd * by_[b];
            e_y += fe * by_[b];
        }
        ox += dx_[a] * e_y;
        oy += bx_[a] * e_dy;
        oz += bx_[a] * e_dz;
        oe += bx_[a] * e_y;
    }
    return make_float4(ox, oy, oz, oe);
}

// the same from a stored stencil sc[(4a+b)*ld]
__device__ __forceinline__ float4 grid_eval_cached(const float4 *__restrict__ sc, int ld, const GridCell &g) {
    return grid_eval_stencil([&](int e) { return sc[e * ld]; }, g);
}

// ---------------------------------------------------------------------------
// reductions (deterministic: xor butterflies give bitwise-identical results
// in every lane; cross-wave partials are summed in wave order)
// ---------------------------------------------------------------------------
// Each step combines a lane's value with one partner's (a + b in one lane,
// b + a in the other: equal bits), so after six steps every lane holds the
// same result.  Partners: xor 1 and xor 2 in a quad (DPP quad_perm), the
// other quad of the half-row and the other half of the row (DPP
// row_half_mirror, row_mirror: the quads and halves already hold equal
// values), the other row of a row pair and the other half of the wave
// (v_permlane16_swap, v_permlane32_swap).  Registers only, no LDS.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    // every lane's source is valid for the patterns used (quad_perm, row
    // mirrors), so bound_ctrl only spares the copy of the unused old value
    const unsigned lo = __builtin_amdgcn_update_dpp(0u, (unsigned)u, CTRL, 0xf, 0xf, true);
    const unsigned hi = __builtin_amdgcn_update_dpp(0u, (unsigned)(u >> 32), CTRL, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <bool ROW16>
__device__ __forceinline__ void swap_halves(double v, double &a, double &b) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
    const auto l = ROW16 ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                         : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = ROW16 ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                         : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    a = __builtin_bit_cast(double, ((unsigned long long)h[0] << 32) | l[0]);
    b = __builtin_bit_cast(double, ((unsigned long long)h[1] << 32) | l[1]);
}
template <typename Op>
__device__ __forceinline__ double wave_reduce(double v, Op op) {
    v = op(v, dpp_f64<0xB1>(v));    // quad_perm [1,0,3,2]
    v = op(v, dpp_f64<0x4E>(v));    // quad_perm [2,3,0,1]
    v = op(v, dpp_f64<0x141>(v));   // row_half_mirror
    v = op(v, dpp_f64<0x140>(v));   // row_mirror
    double a, b;
    swap_halves<true>(v, a, b);     // rows (0,1), (2,3): a = the even row's, b = the odd row's
    v = op(a, b);
    swap_halves<false>(v, a, b);    // lanes 0-31 and 32-63
    return op(a, b);
}
__device__ __forceinline__ double wave_sum(double v) {
    return wave_reduce(v, [](double x, double y) { return x + y; });
}

// Lanes 0-31 of x and 32-63 of y exchanged (ROW16: the odd rows of x and the
// even rows of y), then added: with ROW16 = false, lanes 0-31 hold
// x[l] + x[l+32] and lanes 32-63 y[l-32] + y[l]; with ROW16 = true, rows
// 0 and 2 hold x[r] + x[r+16] and rows 1 and 3 y[r-16] + y[r].
template <bool ROW16>
__device__ __forceinline__ double swap_add(double x, double y) {
    const unsigned long long ux = __builtin_bit_cast(unsigned long long, x);
    const unsigned long long uy = __builtin_bit_cast(unsigned long long, y);
    const auto l = ROW16 ? __builtin_amdgcn_permlane16_swap((unsigned)ux, (unsigned)uy, false, false)
                         : __builtin_amdgcn_permlane32_swap((unsigned)ux, (unsigned)uy, false, false);
    const auto h = ROW16 ? __builtin_amdgcn_permlane16_swap((unsigned)(ux >> 32), (unsigned)(uy >> 32), false, false)
                         : __builtin_amdgcn_permlane32_swap((unsigned)(ux >> 32), (unsigned)(uy >> 32), false, false);
    const double a = __builtin_bit_cast(double, ((unsigned long long)h[0] << 32) | l[0]);
    const double b = __builtin_bit_cast(double, ((unsigned long long)h[1] << 32) | l[1]);
    return a + b;
}

// Wave sums of three values, wave-uniform (scalar) results.  Reduce-scatter:
// the half-wave exchange leaves a's half sums in lanes 0-31 and b's in lanes
// 32-63 (c is summed over the halves in every lane), the row exchange leaves
// a in row 0, c in rows 1 and 3 and b in row 2, and four DPP steps sum each
// row; the results are read from lanes 0, 32 and 16.  Every sum runs in a
// fixed order, the same for every wave.
__device__ __forceinline__ void wave_sum3(double a, double b, double c, double &ra, double &rb, double &rc) {
    const double s = swap_add<false>(a, b);
    const double c1 = swap_add<false>(c, c);
    double t = swap_add<true>(s, c1);
    t += dpp_f64<0xB1>(t);    // quad_perm [1,0,3,2]
    t += dpp_f64<0x4E>(t);    // quad_perm [2,3,0,1]
    t += dpp_f64<0x141>(t);   // row_half_mirror
    t += dpp_f64<0x140>(t);   // row_mirror
    const unsigned long long u = __builtin_bit_cast(unsigned long long, t);
    auto rd = [&](int lane) {
        const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, lane);
        const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
        return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    };
    ra = rd(0);
    rb = rd(32);
    rc = rd(16);
}
__device__ __forceinline__ double wave_max(double v) {
    return wave_reduce(v, [](double x, double y) { return fmax(x, y); });
}
