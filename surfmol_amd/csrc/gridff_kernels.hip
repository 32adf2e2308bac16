// gridff_kernels.hip -- GridFF construction on the device, FP64.
//
//  k_gridff_build   reference Pauli / London / Coulomb grid by a brute sum
//                   over substrate atoms and lateral images
//                   (make_GridFF, opencl/gridff_build.cl:908-1062), one
//                   thread per node, written z-fastest with 4 channels.
//  k_line_solve     one 1-D B-spline interpolation solve per thread:
//                   (1/6, 2/3, 1/6) tridiagonal, truncated (z) or cyclic
//                   (x, y; Sherman-Morrison around a Thomas sweep).  Three
//                   passes give the exact solution of (Sx (x) Sy (x) Sz) C = E,
//                   the fixed point of the reference's iterative fit driver
//                   around BsplineConv3D (opencl/gridff_build.cl:8-61).
//  k_pack_f32       FP64 -> FP32 coefficients (float4 per node).
#include "smr_kernels.h"

namespace smr {

__global__ void k_gridff_build(GridBuildArgs G) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ntot = (size_t)G.nx * G.ny * G.nz;
    if (n >= ntot) return;
    const int iz = (int)(n % G.nz);
    const int iy = (int)((n / G.nz) % G.ny);
    const int ix = (int)(n / ((size_t)G.nz * G.ny));
    // node position: p0 + lvec_a*ia/nx + lvec_b*ib/ny + lvec_c*ic/nz (gridff_build.cl:1006-1007)
    const double px = G.p0[0] + G.L[0] * ix / (double)G.nx;
    const double py = G.p0[1] + G.L[1] * iy / (double)G.ny;
    const double pz = G.p0[2] + G.L[2] * iz / (double)G.nz;
    double eP = 0.0, eL = 0.0, eC = 0.0;
    for (int ja = 0; ja < G.nsub; ja++) {
        const double ax = G.sub_pos[3 * ja], ay = G.sub_pos[3 * ja + 1], az = G.sub_pos[3 * ja + 2];
        const double R = G.sub_req[4 * ja], sE = G.sub_req[4 * ja + 1], Q = G.sub_req[4 * ja + 2];
        for (int jy = -G.npbc; jy <= G.npbc; jy++) {
            for (int jx = -G.npbc; jx <= G.npbc; jx++) {
                const double dx = px - ax + G.L[0] * jx;
                const double dy = py - ay + G.L[1] * jy;
                const double dz = pz - az;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double r = sqrt(r2 + 1e-32);
                const double ir2 = 1.0 / (r2 + G.r2damp);
                eC += 14.3996448915 * Q * sqrt(ir2);
                const double e = exp(-G.alpha * (r - R));
                const double eM = sE * e;
                eP += eM * e;
                eL += -2.0 * eM;
            }
        }
    }
    double *o = G.out + 4 * n;
    o[0] = eP;
    o[1] = eL;
    o[2] = eC;
    o[3] = 0.0;
}

struct LineSolve {
    double *data;
    int n;                 // line length
    int n2;                // lines = n1 * n2; line l -> (l / n2, l % n2)
    long long nlines;
    long long s1, s2, se;  // strides (in doubles) of l1, l2 and along the line
    const double *cp;      // Thomas c'_i
    const double *inv_den; // 1 / (b_i - a c'_{i-1})
    const double *zvec;    // Sherman-Morrison z = T^-1 u (cyclic only)
    double vlast;          // v_{n-1} = beta/gamma
    double fac;            // 1 / (1 + v.z)
    int cyclic;
};

__global__ void k_line_solve(LineSolve S) {
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= S.nlines) return;
    const long long l1 = l / S.n2, l2 = l % S.n2;
    double *base = S.data + l1 * S.s1 + l2 * S.s2;
    const double a = 1.0 / 6.0;
    // forward sweep (4 channels)
    double d0[4];
    {
        double *p = base;
        for (int c = 0; c < 4; c++) { d0[c] = p[c] * S.inv_den[0]; p[c] = d0[c]; }
    }
    for (int i = 1; i < S.n; i++) {
        double *p = base + i * S.se;
        const double id = S.inv_den[i];
        for (int c = 0; c < 4; c++) { d0[c] = (p[c] - a * d0[c]) * id; p[c] = d0[c]; }
    }
    // back substitution
    double x[4] = {d0[0], d0[1], d0[2], d0[3]};
    for (int i = S.n - 2; i >= 0; i--) {
        double *p = base + i * S.se;
        const double cpi = S.cp[i];
        for (int c = 0; c < 4; c++) { x[c] = p[c] - cpi * x[c]; p[c] = x[c]; }
    }
    if (S.cyclic) {
        const double *pN = base + (long long)(S.n - 1) * S.se;
        double f[4];
        for (int c = 0; c < 4; c++) f[c] = (base[c] + S.vlast * pN[c]) * S.fac;
        for (int i = 0; i < S.n; i++) {
            double *p = base + i * S.se;
            const double zi = S.zvec[i];
            for (int c = 0; c < 4; c++) p[c] -= f[c] * zi;
        }
    }
}

__global__ void k_pack_f32(const double *C, float4 *out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = make_float4((float)C[4 * i], (float)C[4 * i + 1], (float)C[4 * i + 2], (float)C[4 * i + 3]);
}

hipError_t launch_gridff_build(const GridBuildArgs &G, hipStream_t st) {
    const size_t ntot = (size_t)G.nx * G.ny * G.nz;
    const int B = 256;
    hipLaunchKernelGGL(k_gridff_build, dim3((unsigned)((ntot + B - 1) / B)), dim3(B), 0, st, G);
    return hipGetLastError();
}

// Host-side coefficients of the 1-D solves (FP64).  Tridiagonal T with
// a = c = 1/6, diagonal 2/3; for the cyclic case the corners are folded in by
// Sherman-Morrison with gamma = -b: A = T' + u v^T, u = (gamma,0..0,alpha),
// v = (1,0..0,beta/gamma), T'_00 = b - gamma, T'_nn = b - alpha*beta/gamma.
struct LineCoefs {
    std::vector<double> cp, inv_den, z;
    double vlast = 0, fac = 0;
};

static void thomas_coefs(int n, const std::vector<double> &diag, std::vector<double> &cp, std::vector<double> &inv_den) {
    const double a = 1.0 / 6.0, c = 1.0 / 6.0;
    cp.assign(n, 0.0);
    inv_den.assign(n, 0.0);
    double cprev = 0.0;
    for (int i = 0; i < n; i++) {
        double den = diag[i] - (i > 0 ? a * cprev : 0.0);
        inv_den[i] = 1.0 / den;
        cp[i] = c * inv_den[i];
        cprev = cp[i];
    }
}

static LineCoefs line_coefs(int n, bool cyclic) {
    const double b = 2.0 / 3.0, alpha = 1.0 / 6.0, beta = 1.0 / 6.0;
    LineCoefs L;
    std::vector<double> diag(n, b);
    if (!cyclic) {
        thomas_coefs(n, diag, L.cp, L.inv_den);
        return L;
    }
    const double gamma = -b;
    diag[0] = b - gamma;
    diag[n - 1] = b - alpha * beta / gamma;
    thomas_coefs(n, diag, L.cp, L.inv_den);
    // solve T' z = u on the host
    std::vector<double> u(n, 0.0), d(n);
    u[0] = gamma;
    u[n - 1] = alpha;
    const double a = 1.0 / 6.0;
    d[0] = u[0] * L.inv_den[0];
    for (int i = 1; i < n; i++) d[i] = (u[i] - a * d[i - 1]) * L.inv_den[i];
    L.z.assign(n, 0.0);
    L.z[n - 1] = d[n - 1];
    for (int i = n - 2; i >= 0; i--) L.z[i] = d[i] - L.cp[i] * L.z[i + 1];
    L.vlast = beta / gamma;
    L.fac = 1.0 / (1.0 + L.z[0] + L.vlast * L.z[n - 1]);
    return L;
}

static hipError_t solve_axis(double *E, int n, long long n1, long long n2, long long s1, long long s2, long long se,
                             bool cyclic, hipStream_t st) {
    LineCoefs C = line_coefs(n, cyclic);
    double *dcoef = nullptr;
    const size_t nb = sizeof(double) * (size_t)n * 3;
    hipError_t e = hipMalloc(&dcoef, nb);
    if (e != hipSuccess) return e;
    std::vector<double> h(3 * (size_t)n, 0.0);
    std::copy(C.cp.begin(), C.cp.end(), h.begin());
    std::copy(C.inv_den.begin(), C.inv_den.end(), h.begin() + n);
    if (cyclic) std::copy(C.z.begin(), C.z.end(), h.begin() + 2 * n);
    e = hipMemcpyAsync(dcoef, h.data(), nb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        LineSolve S;
        S.data = E;
        S.n = n;
        S.n2 = (int)n2;
        S.nlines = n1 * n2;
        S.s1 = s1; S.s2 = s2; S.se = se;
        S.cp = dcoef;
        S.inv_den = dcoef + n;
        S.zvec = dcoef + 2 * n;
        S.vlast = C.vlast;
        S.fac = C.fac;
        S.cyclic = cyclic ? 1 : 0;
        const int B = 128;
        hipLaunchKernelGGL(k_line_solve, dim3((unsigned)((S.nlines + B - 1) / B)), dim3(B), 0, st, S);
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamSynchronize(st);
    (void)hipFree(dcoef);
    return e != hipSuccess ? e : e2;
}

hipError_t launch_bspline_fit(double *E, int nx, int ny, int nz, double * /*work*/, hipStream_t st) {
    const long long NZ4 = 4LL * nz, NYZ4 = 4LL * ny * nz;
    hipError_t e;
    // z lines: lines (ix, iy), element stride 4
    if ((e = solve_axis(E, nz, nx, ny, NYZ4, NZ4, 4, false, st)) != hipSuccess) return e;
    // y lines: lines (ix, iz), element stride 4*nz
    if ((e = solve_axis(E, ny, nx, nz, NYZ4, 4, NZ4, true, st)) != hipSuccess) return e;
    // x lines: lines (iy, iz), element stride 4*ny*nz
    if ((e = solve_axis(E, nx, ny, nz, NZ4, 4, NYZ4, true, st)) != hipSuccess) return e;
    return hipSuccess;
}

hipError_t launch_pack_f32(const double *C, float *out, size_t n4, hipStream_t st) {
    const int B = 256;
    hipLaunchKernelGGL(k_pack_f32, dim3((unsigned)((n4 + B - 1) / B)), dim3(B), 0, st, C,
                       reinterpret_cast<float4 *>(out), n4);
    return hipGetLastError();
}

}  // namespace smr
