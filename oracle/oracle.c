/*
 * oracle.c -- FP64 CPU restatement of SurfMol's multi-system relaxation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library under
 * surfmol_amd/) links, loads or calls this file; only tests/, the smoke check
 * and bench.py's cpu_baseline leg use it, as the checker and as the timed
 * single-thread CPU baseline ("port").
 *
 * Every function restates one reference function; the citation is given at
 * each one (paths relative to the reference repository root).  Arithmetic is
 * double precision throughout and follows the reference's order of
 * operations, so that results agree with the reference's own CPU path to
 * rounding.
 *
 * Parity status of each part (how it is pinned) is listed in the header of
 * oracle/oracle.py and in tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_COULOMB_CONST 14.3996448915 /* eV*A/e^2, molff/src/nonbonded.rs:8 */
#define ORC_EXCL_MAX 16                 /* molff/src/nonbonded.rs:7 */

/* Convention switches for the non-bonded pair term.  The reference CPU
 * (NonBondedFF) and GPU (getNonBond) paths disagree; both are restated.     */
enum {
    ORC_NB_OFF = 0,
    ORC_NB_GPU = 1, /* opencl/UFF.cl:57-95,1023-1204: E0=REQi.y*REQj.y, 1-2 exclusion via neighs, no H term, no clamp */
    ORC_NB_CPU = 2  /* molff/src/nonbonded.rs:166-243: E0=(sqrtEi*sqrtEj)^2, 1-2+1-3 excl list, H term, clamp, rcut */
};

typedef struct {
    /* ---- topology (all indices 0-based, local to the system) ---- */
    int natoms, nbonds, nangles, ndih, ninv;
    const int *neighs;        /* [4*natoms]  moltopo/src/molecular.rs:34-78   */
    const int *neigh_bs;      /* [4*natoms]  bond index per neighbour slot    */
    const double *bon_params; /* [2*nbonds]  (k, l0)       uff.rs:77          */
    const int *ang_atoms;     /* [3*nangles] (i, j, k), j central              */
    const int *ang_ngs;       /* [2*nangles] hneigh slots  uff.rs:250-267     */
    const double *ang_params; /* [5*nangles] (k,c0,c1,c2,c3) uff.rs:78        */
    const int *dih_atoms;     /* [4*ndih]                                      */
    const int *dih_ngs;       /* [3*ndih]                  uff.rs:269-294     */
    const double *dih_params; /* [3*ndih]    (V, d, n)     uff.rs:79          */
    const int *inv_atoms;     /* [4*ninv]    i central                         */
    const int *inv_ngs;       /* [3*ninv]                  uff.rs:296-315     */
    const double *inv_params; /* [4*ninv]    (K, C0, C1, C2) uff.rs:80        */
    /* ---- non-bonded ---- */
    int nb_mode;              /* ORC_NB_* */
    const double *reqs;       /* [4*natoms] (RvdW, sqrt(EvdW), Q, Hb) params.rs:595-604 */
    const int *excl;          /* [16*natoms] sorted, -1 padded (CPU convention only)    */
    double rdamp;             /* damping radius; R2damp = rdamp^2                       */
    double fmax_nb;           /* CPU convention clamp (nonbonded.rs:86)                 */
    int clamp_nb;             /* CPU convention: apply clamp                            */
    double rcut2;             /* CPU convention: hard cutoff squared (1e300 = none)     */
    /* ---- GridFF (tricubic B-spline, z-fastest (P,L,Q,pad) coefficients) ---- */
    const double *grid;       /* [nx*ny*nz*4] or NULL */
    int gnx, gny, gnz;
    double gp0[3];
    double ginv_step[3];
    double alpha_morse;       /* GFFParams.y (UFF.cl:1681) */
    /* coupled layer of molecules (GPU convention only): pairs of atoms of
     * different molecules interact only within the hard cutoff, r^2 < rcut_inter2
     * (NonBondedFF::eval_broad culls pairs beyond rc, nonbonded.rs:251-324);
     * pairs inside one molecule always interact. NULL = one molecule. */
    const int *mol_id;        /* [natoms] */
    double rcut_inter2;
    int layer_cells;          /* 1: find intermolecular pairs through a cell grid (same terms) */
    double rswitch_inter;     /* > 0: intermolecular pair energies are switched off smoothly over
                                 [rc - rswitch_inter, rc] by S(x) = 1 - 3x^2 + 2x^3; 0: hard cut */
    /* coupled layer with lateral periodic images (the 3x3 image loop of
     * getNonBond's PBC branch, opencl/UFF.cl:1154-1182, and NonBondedFF::
     * eval_pbc, nonbonded.rs:376-418, with a cutoff): box lengths along x and
     * y (0: open).  Every pair of atoms, in every image within the cutoff,
     * interacts as an intermolecular pair -- except an atom with the same
     * image of its own molecule, which is the intramolecular pair. */
    double pbc_lx, pbc_ly;
    /* optional [natoms]: each force contribution's magnitude is added to its
     * atom's entry (the scale of the per-atom force sum; NULL = off) */
    double *fscale;
} orc_system;


typedef struct {
    double dt0, dt_max;       /* FireState::new(dt0, dt_max) raff.rs:1057-1060 */
    double alpha0;            /* 0.1  */
    int n_min;                /* 5    */
    double f_inc, f_dec, f_alpha; /* 1.1, 0.5, 0.99 */
    double flim;              /* <=0: no clamp (relax_pentacene_uff.rs loop); >0: clamp as raff.rs:1093 */
} orc_fire_params;

/* ------------------------------------------------------------------ */
/* small vector helpers                                                */
/* ------------------------------------------------------------------ */
typedef struct { double x, y, z; } v3;
static inline v3 v3_ld(const double *p) { v3 r = {p[0], p[1], p[2]}; return r; }
static inline void v3_st(double *p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline v3 v3_sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline v3 v3_add(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline v3 v3_mul(v3 a, double s) { v3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static inline double v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 v3_cross(v3 a, v3 b) {
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
/* Vec3d::set_lincomb(a, u, b, w) = a*u + b*w */
static inline v3 v3_lincomb(double a, v3 u, double b, v3 w) {
    v3 r = {a * u.x + b * w.x, a * u.y + b * w.y, a * u.z + b * w.z};
    return r;
}
/* per-atom force scale: adds |f| of one contribution to atom ia's entry */
#define ORC_SCALE(s, ia, f) do { if ((s)->fscale) (s)->fscale[(ia)] += sqrt(v3_dot((f), (f))); } while (0)

/* Vec2d::mul_cmplx */
static inline void cmul(double *x, double *y, double cx, double cy) {
    double nx = (*x) * cx - (*y) * cy;
    double ny = (*x) * cy + (*y) * cx;
    *x = nx; *y = ny;
}

/* ------------------------------------------------------------------ */
/* UFF bonded terms                                                    */
/* ------------------------------------------------------------------ */

/* Uff::eval_atom_bonds  (molff/src/uff.rs:440-468).
 * hneigh[4*(4*ia+k)] = (h.x, h.y, h.z, 1/l); adds bond force to fapos[ia];
 * returns sum over neighbours of k*dl^2 (bond energy counted at both ends). */
static double orc_eval_atom_bonds(const orc_system *s, int ia, const double *apos, double *fapos, double *hneigh) {
    v3 pa = v3_ld(apos + 3 * ia);
    double e = 0.0;
    for (int in = 0; in < 4; in++) {
        int ing = s->neighs[4 * ia + in];
        if (ing < 0) break;
        int inn = ia * 4 + in;
        v3 pb = v3_ld(apos + 3 * ing);
        v3 dp = v3_sub(pb, pa);
        double l = sqrt(v3_dot(dp, dp));
        double invl = 1.0 / l;
        dp = v3_mul(dp, invl);
        hneigh[4 * inn + 0] = dp.x;
        hneigh[4 * inn + 1] = dp.y;
        hneigh[4 * inn + 2] = dp.z;
        hneigh[4 * inn + 3] = invl;
        int ib = s->neigh_bs[4 * ia + in];
        double k = s->bon_params[2 * ib + 0];
        double l0 = s->bon_params[2 * ib + 1];
        double dl = l - l0;
        e += k * dl * dl;
        double fr = 2.0 * k * dl;
        v3 f = v3_mul(dp, fr);
        v3 fa = v3_ld(fapos + 3 * ia);
        v3_st(fapos + 3 * ia, v3_add(fa, f));
        ORC_SCALE(s, ia, f);
    }
    return e;
}

/* Uff::eval_angle_prokop  (molff/src/uff.rs:471-507).  Pieces (fpi,fpj,fpk)
 * written to fint[i0ang + 3*ia + {0,1,2}].  The energy omits k*c1*cos(theta)
 * exactly as the reference does (the force keeps c1). */
static double orc_eval_angle(const orc_system *s, int ia, const double *hneigh, double *fint, int i0ang) {
    int g0 = s->ang_ngs[2 * ia + 0], g1 = s->ang_ngs[2 * ia + 1];
    if (g0 < 0 || g1 < 0) return 0.0;
    v3 qij = v3_ld(hneigh + 4 * g0); double qij_w = hneigh[4 * g0 + 3];
    v3 qkj = v3_ld(hneigh + 4 * g1); double qkj_w = hneigh[4 * g1 + 3];
    v3 h = v3_add(qij, qkj);
    double c = 0.5 * (v3_dot(h, h) - 2.0);
    double sn = sqrt(1.0 - c * c + 1e-14);
    double inv_sin = 1.0 / sn;
    const double *par = s->ang_params + 5 * ia;
    double ee = par[1];
    double ff = par[2];
    double cx = c, cy = sn;
    cmul(&cx, &cy, c, sn); /* (cos 2t, sin 2t) */
    ee += par[3] * cx;
    ff += par[3] * cy * inv_sin * 2.0;
    cmul(&cx, &cy, c, sn); /* (cos 3t, sin 3t) */
    ee += par[4] * cx;
    ff += par[4] * cy * inv_sin * 3.0;
    ee *= par[0];
    ff *= par[0];
    double fi = ff * qij_w;
    double fk = ff * qkj_w;
    double fic = fi * c;
    double fkc = fk * c;
    v3 fpi = v3_lincomb(fic, qij, -fi, qkj);
    v3 fpk = v3_lincomb(-fk, qij, fkc, qkj);
    v3 fpj = v3_lincomb(fk - fic, qij, fi - fkc, qkj);
    int i0 = ia * 3 + i0ang;
    v3_st(fint + 3 * (i0 + 0), fpi);
    v3_st(fint + 3 * (i0 + 1), fpj);
    v3_st(fint + 3 * (i0 + 2), fpk);
    return ee;
}

/* Uff::eval_dihedral_prokop  (molff/src/uff.rs:510-545). */
static double orc_eval_dihedral(const orc_system *s, int id, const double *hneigh, double *fint, int i0dih) {
    int g0 = s->dih_ngs[3 * id + 0], g1 = s->dih_ngs[3 * id + 1], g2 = s->dih_ngs[3 * id + 2];
    if (g0 < 0 || g1 < 0 || g2 < 0) return 0.0;
    v3 q12 = v3_ld(hneigh + 4 * g0); double q12w = hneigh[4 * g0 + 3];
    v3 q32 = v3_ld(hneigh + 4 * g1); double q32w = hneigh[4 * g1 + 3];
    v3 q43 = v3_ld(hneigh + 4 * g2); double q43w = hneigh[4 * g2 + 3];
    v3 n123 = v3_cross(q12, q32);
    v3 n234 = v3_cross(q43, q32);
    double il2_123 = 1.0 / v3_dot(n123, n123);
    double il2_234 = 1.0 / v3_dot(n234, n234);
    double inv_n12 = sqrt(il2_123 * il2_234);
    double csx = v3_dot(n123, n234) * inv_n12;
    double csy = -v3_dot(n123, q43) * inv_n12;
    double cnx = csx, cny = csy;
    const double *par = s->dih_params + 3 * id;
    int n = (int)par[2];
    for (int i = 1; i < n; i++) cmul(&cnx, &cny, csx, csy);
    double ee = par[0] * (1.0 + par[1] * cnx);
    double f = -par[0] * par[1] * par[2] * cny;
    v3 fp1 = v3_mul(n123, -f * il2_123 * q12w);
    v3 fp4 = v3_mul(n234, f * il2_234 * q43w);
    double c123 = v3_dot(q32, q12) * (q32w / q12w);
    double c432 = v3_dot(q32, q43) * (q32w / q43w);
    v3 fp3 = v3_lincomb(-c123, fp1, -c432 - 1.0, fp4);
    v3 fp2 = v3_lincomb(c123 - 1.0, fp1, c432, fp4);
    int i0 = id * 4 + i0dih;
    v3_st(fint + 3 * (i0 + 0), fp1);
    v3_st(fint + 3 * (i0 + 1), fp2);
    v3_st(fint + 3 * (i0 + 2), fp3);
    v3_st(fint + 3 * (i0 + 3), fp4);
    return ee;
}

/* Uff::eval_inversion_prokop  (molff/src/uff.rs:548-579). */
static double orc_eval_inversion(const orc_system *s, int ii, const double *hneigh, double *fint, int i0inv) {
    int g0 = s->inv_ngs[3 * ii + 0], g1 = s->inv_ngs[3 * ii + 1], g2 = s->inv_ngs[3 * ii + 2];
    if (g0 < 0 || g1 < 0 || g2 < 0) return 0.0;
    v3 q21 = v3_ld(hneigh + 4 * g0); double q21w = hneigh[4 * g0 + 3];
    v3 q31 = v3_ld(hneigh + 4 * g1); double q31w = hneigh[4 * g1 + 3];
    v3 q41 = v3_ld(hneigh + 4 * g2); double q41w = hneigh[4 * g2 + 3];
    v3 n123 = v3_cross(q21, q31);
    double nl = sqrt(v3_dot(n123, n123)); /* Vec3d::normalize returns the old norm */
    n123 = v3_mul(n123, 1.0 / nl);
    double il123 = 1.0 / nl;
    double sn = -v3_dot(n123, q41);
    double c = sqrt(1.0 - sn * sn + 1e-14);
    const double *par = s->inv_params + 4 * ii;
    double c2x = c, c2y = sn;
    cmul(&c2x, &c2y, c, sn);
    double ee = par[0] * (par[1] + par[2] * c + par[3] * c2x);
    double f = -par[0] * (par[2] * sn + 2.0 * par[3] * c2y) / c;
    double fq41 = f * q41w;
    double fi123 = f * il123;
    v3 tq = v3_lincomb(sn * fi123, n123, fi123, q41);
    v3 fp4 = v3_lincomb(fq41, n123, sn * fq41, q41);
    v3 fp2 = v3_mul(v3_cross(q31, tq), q21w);
    v3 fp3 = v3_mul(v3_cross(tq, q21), q31w);
    v3 fp1 = v3_mul(v3_add(v3_add(fp2, fp3), fp4), -1.0);
    int i0 = ii * 4 + i0inv;
    v3_st(fint + 3 * (i0 + 0), fp1);
    v3_st(fint + 3 * (i0 + 1), fp2);
    v3_st(fint + 3 * (i0 + 2), fp3);
    v3_st(fint + 3 * (i0 + 3), fp4);
    return ee;
}

/* Uff::map_atom_interactions + assemble_atom_force (molff/src/uff.rs:317-386):
 * each atom gathers its pieces in the order dihedrals, inversions, angles
 * (the bucket fill order), adding them onto the bond force already in fapos. */
static void orc_assemble(const orc_system *s, const double *fint, double *fapos) {
    int na = s->natoms;
    int i0dih = 0, i0inv = 4 * s->ndih, i0ang = i0inv + 4 * s->ninv;
    int *cnt = (int *)calloc((size_t)na, sizeof(int));
    int *off = (int *)calloc((size_t)na + 1, sizeof(int));
    int nobj = i0ang + 3 * s->nangles;
    int *obj = (int *)malloc(sizeof(int) * (size_t)(nobj > 0 ? nobj : 1));
    for (int i = 0; i < s->ndih; i++) for (int k = 0; k < 4; k++) cnt[s->dih_atoms[4 * i + k]]++;
    for (int i = 0; i < s->ninv; i++) for (int k = 0; k < 4; k++) cnt[s->inv_atoms[4 * i + k]]++;
    for (int i = 0; i < s->nangles; i++) for (int k = 0; k < 3; k++) cnt[s->ang_atoms[3 * i + k]]++;
    for (int i = 0; i < na; i++) off[i + 1] = off[i] + cnt[i];
    memset(cnt, 0, sizeof(int) * (size_t)na);
    for (int i = 0; i < s->ndih; i++) for (int k = 0; k < 4; k++) { int a = s->dih_atoms[4 * i + k]; obj[off[a] + cnt[a]++] = i0dih + 4 * i + k; }
    for (int i = 0; i < s->ninv; i++) for (int k = 0; k < 4; k++) { int a = s->inv_atoms[4 * i + k]; obj[off[a] + cnt[a]++] = i0inv + 4 * i + k; }
    for (int i = 0; i < s->nangles; i++) for (int k = 0; k < 3; k++) { int a = s->ang_atoms[3 * i + k]; obj[off[a] + cnt[a]++] = i0ang + 3 * i + k; }
    for (int ia = 0; ia < na; ia++) {
        v3 f = v3_ld(fapos + 3 * ia);
        for (int i = off[ia]; i < off[ia + 1]; i++) {
            v3 fp = v3_ld(fint + 3 * obj[i]);
            f = v3_add(f, fp);
            ORC_SCALE(s, ia, fp);
        }
        v3_st(fapos + 3 * ia, f);
    }
    free(cnt); free(off); free(obj);
}

/* Uff::eval_forces  (molff/src/uff.rs:611-637).  Zeroes fapos, returns
 * (eb, ea, ed, ei) in E4.  hneigh/fint may be NULL (scratch allocated). */
void orc_uff_eval(const orc_system *s, const double *apos, double *fapos, double *E4) {
    int na = s->natoms;
    int nf = 4 * s->ndih + 4 * s->ninv + 3 * s->nangles;
    double *hneigh = (double *)calloc((size_t)na * 16, sizeof(double));
    double *fint = (double *)calloc((size_t)(nf > 0 ? nf : 1) * 3, sizeof(double));
    double eb = 0, ea = 0, ed = 0, ei = 0;
    for (int i = 0; i < 3 * na; i++) fapos[i] = 0.0;
    for (int ia = 0; ia < na; ia++) eb += orc_eval_atom_bonds(s, ia, apos, fapos, hneigh);
    int i0dih = 0, i0inv = 4 * s->ndih, i0ang = i0inv + 4 * s->ninv;
    for (int i = 0; i < s->nangles; i++) ea += orc_eval_angle(s, i, hneigh, fint, i0ang);
    for (int i = 0; i < s->ndih; i++) ed += orc_eval_dihedral(s, i, hneigh, fint, i0dih);
    for (int i = 0; i < s->ninv; i++) ei += orc_eval_inversion(s, i, hneigh, fint, i0inv);
    orc_assemble(s, fint, fapos);
    E4[0] = eb; E4[1] = ea; E4[2] = ed; E4[3] = ei;
    free(hneigh); free(fint);
}

/* ------------------------------------------------------------------ */
/* Non-bonded pair term                                                */
/* ------------------------------------------------------------------ */

/* getLJQH (opencl/UFF.cl:82-95) in FP64: REQ already mixed.  Returns E and
 * writes the force on atom i (dp = pj - pi) into *f.                        */
static inline double orc_ljqh_gpu(v3 dp, double R0, double E0, double Q, double r2damp, v3 *f) {
    double r2 = v3_dot(dp, dp);
    double ir2_ = 1.0 / (r2 + r2damp);
    double Ec = ORC_COULOMB_CONST * Q * sqrt(ir2_);
    double ir2 = 1.0 / r2;
    double u2 = R0 * R0 * ir2;
    double u6 = u2 * u2 * u2;
    double vdW = u6 * E0;
    double E = (u6 - 2.0) * vdW + Ec;
    double fr = -12.0 * (u6 - 1.0) * vdW * ir2 - Ec * ir2_;
    *f = v3_mul(dp, fr);
    return E;
}

/* NonBondedFF::get_ljqh (molff/src/nonbonded.rs:171-186). */
static inline double orc_ljqh_cpu(v3 dp, const double *reqh, double r2damp, v3 *f) {
    double r2 = v3_dot(dp, dp);
    double ir2_ = 1.0 / (r2 + r2damp);
    double e_coul = ORC_COULOMB_CONST * reqh[2] * sqrt(ir2_);
    double f_coul = e_coul * ir2_;
    double ir2 = 1.0 / r2;
    double u2 = reqh[0] * reqh[0] * ir2;
    double u6 = u2 * u2 * u2;
    double e_vdw = reqh[1] * reqh[1];
    double vdw = u6 * e_vdw;
    double h = u6 * u6 * (reqh[3] < 0.0 ? reqh[3] * e_vdw : 0.0);
    double e = e_coul + (u6 - 2.0) * vdw + h;
    double fm = f_coul + ((u6 - 1.0) * vdw + h) * ir2 * 12.0;
    *f = v3_mul(dp, -fm);
    return e;
}

/* Non-bonded pairs over all ordered pairs (each pair visited from both ends,
 * energy therefore double counted, as in both reference paths).
 *   GPU convention: getNonBond (opencl/UFF.cl:1023-1204), non-PBC branch.
 *   CPU convention: NonBondedFF::eval / eval_nb_atom (nonbonded.rs:199-243).
 * Adds forces into fapos, returns the energy.                               */
/* Intermolecular pair of a coupled layer within the cutoff: getLJQH, times
 * the switch S(r) when a switching width is set (E S; force (fr S + E S'/r) dp). */
static inline double orc_inter_pair(const orc_system *s, v3 dp, const double *ri, const double *rj, double r2damp,
                                    v3 *f) {
    const double e = orc_ljqh_gpu(dp, ri[0] + rj[0], ri[1] * rj[1], ri[2] * rj[2], r2damp, f);
    if (s->rswitch_inter <= 0.0) return e;
    const double rc = sqrt(s->rcut_inter2), ron = rc - s->rswitch_inter, r2 = v3_dot(dp, dp);
    if (r2 <= ron * ron) return e;
    const double r = sqrt(r2), w = s->rswitch_inter, x = (r - ron) / w;
    const double S = 1.0 - x * x * (3.0 - 2.0 * x), dS_r = 6.0 * x * (x - 1.0) / w / r;
    *f = v3_add(v3_mul(*f, S), v3_mul(dp, e * dS_r));
    return e * S;
}

/* Coupled layer, GPU convention: intramolecular pairs as getNonBond (all
 * pairs of the molecule but 1-2 neighbours), intermolecular pairs within the
 * cutoff found through a uniform cell grid of cell size >= rc (the broad
 * phase of NonBondedFF::eval_broad, nonbonded.rs:251-324, with the cell grid
 * of spacc/src/uniform_grid.rs:87-116).  Molecules are contiguous atom
 * ranges (mol_id non-decreasing).  Same terms as the brute-force loop, in
 * another summation order.                                                 */
static double orc_nb_eval_layer(const orc_system *s, const double *apos, double *fapos) {
    const int na = s->natoms;
    const double r2damp = s->rdamp * s->rdamp, rc2 = s->rcut_inter2, rc = sqrt(rc2);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < na; i++)
        for (int k = 0; k < 3; k++) {
            if (apos[3 * i + k] < lo[k]) lo[k] = apos[3 * i + k];
            if (apos[3 * i + k] > hi[k]) hi[k] = apos[3 * i + k];
        }
    double h = rc;
    int n[3];
    long long ncell;
    for (;;) {
        ncell = 1;
        for (int k = 0; k < 3; k++) {
            n[k] = (int)((hi[k] - lo[k]) / h) + 1;
            ncell *= n[k];
        }
        if (ncell <= (1 << 22)) break;
        h *= 2.0;
    }
    int *cell = (int *)malloc(sizeof(int) * (size_t)na), *start = (int *)calloc((size_t)ncell + 1, sizeof(int));
    int *order = (int *)malloc(sizeof(int) * (size_t)na), *m0 = (int *)malloc(sizeof(int) * (size_t)na);
    int *m1 = (int *)malloc(sizeof(int) * (size_t)na);
    for (int i = 0; i < na; i++) {
        int c[3];
        for (int k = 0; k < 3; k++) c[k] = (int)((apos[3 * i + k] - lo[k]) / h);
        cell[i] = (c[2] * n[1] + c[1]) * n[0] + c[0];
        start[cell[i] + 1]++;
    }
    for (long long q = 0; q < ncell; q++) start[q + 1] += start[q];
    {
        int *fill = (int *)malloc(sizeof(int) * (size_t)ncell);
        memcpy(fill, start, sizeof(int) * (size_t)ncell);
        for (int i = 0; i < na; i++) order[fill[cell[i]]++] = i;
        free(fill);
    }
    /* atom ranges of the molecules */
    for (int i = 0; i < na;) {
        int j = i;
        while (j < na && s->mol_id[j] == s->mol_id[i]) j++;
        for (int q = i; q < j; q++) { m0[q] = i; m1[q] = j; }
        i = j;
    }
    double etot = 0.0;
    for (int ia = 0; ia < na; ia++) {
        v3 pi = v3_ld(apos + 3 * ia);
        const double *ri = s->reqs + 4 * ia;
        const int *ng = s->neighs + 4 * ia;
        v3 fe = {0, 0, 0};
        double e = 0.0;
        for (int ja = m0[ia]; ja < m1[ia]; ja++) {
            if (ja == ia || ja == ng[0] || ja == ng[1] || ja == ng[2] || ja == ng[3]) continue;
            const double *rj = s->reqs + 4 * ja;
            v3 f;
            e += orc_ljqh_gpu(v3_sub(v3_ld(apos + 3 * ja), pi), ri[0] + rj[0], ri[1] * rj[1], ri[2] * rj[2], r2damp, &f);
            fe = v3_add(fe, f);
            ORC_SCALE(s, ia, f);
        }
        const int cx = cell[ia] % n[0], cy = (cell[ia] / n[0]) % n[1], cz = cell[ia] / (n[0] * n[1]);
        for (int z = cz - 1; z <= cz + 1; z++) {
            if (z < 0 || z >= n[2]) continue;
            for (int y = cy - 1; y <= cy + 1; y++) {
                if (y < 0 || y >= n[1]) continue;
                for (int x = cx - 1; x <= cx + 1; x++) {
                    if (x < 0 || x >= n[0]) continue;
                    const int cc = (z * n[1] + y) * n[0] + x;
                    for (int q = start[cc]; q < start[cc + 1]; q++) {
                        const int ja = order[q];
                        if (s->mol_id[ja] == s->mol_id[ia]) continue;
                        v3 dp = v3_sub(v3_ld(apos + 3 * ja), pi);
                        if (v3_dot(dp, dp) >= rc2) continue;
                        v3 f;
                        e += orc_inter_pair(s, dp, ri, s->reqs + 4 * ja, r2damp, &f);
                        fe = v3_add(fe, f);
                        ORC_SCALE(s, ia, f);
                    }
                }
            }
        }
        etot += e;
        v3_st(fapos + 3 * ia, v3_add(v3_ld(fapos + 3 * ia), fe));
    }
    free(cell); free(start); free(order); free(m0); free(m1);
    return etot;
}

double orc_nb_eval(const orc_system *s, const double *apos, double *fapos) {
    int na = s->natoms;
    double r2damp = s->rdamp * s->rdamp;
    double etot = 0.0;
    const int pbc = s->mol_id && (s->pbc_lx > 0.0 || s->pbc_ly > 0.0);
    if (s->nb_mode == ORC_NB_GPU && s->mol_id && s->layer_cells && !pbc) return orc_nb_eval_layer(s, apos, fapos);
    if (s->nb_mode == ORC_NB_GPU && pbc) {
        /* brute force over atom pairs and the images around the minimum
         * image (a box of at least two cutoffs puts at most one image of a
         * pair within the cutoff) */
        const double L[2] = {s->pbc_lx, s->pbc_ly};
        for (int ia = 0; ia < na; ia++) {
            v3 pi = v3_ld(apos + 3 * ia);
            const double *ri = s->reqs + 4 * ia;
            const int *ng = s->neighs + 4 * ia;
            v3 fe = {0, 0, 0};
            double e = 0.0;
            for (int ja = 0; ja < na; ja++) {
                const double *rj = s->reqs + 4 * ja;
                const v3 dp0 = v3_sub(v3_ld(apos + 3 * ja), pi);
                const int same = s->mol_id[ja] == s->mol_id[ia];
                int m[2] = {0, 0};
                const double d0[2] = {dp0.x, dp0.y};
                for (int k = 0; k < 2; k++)
                    if (L[k] > 0.0) m[k] = -(int)floor(d0[k] / L[k] + 0.5);
                for (int iy = -1; iy <= 1; iy++) {
                    if (L[1] <= 0.0 && iy != 0) continue;
                    for (int ix = -1; ix <= 1; ix++) {
                        if (L[0] <= 0.0 && ix != 0) continue;
                        const int nx = m[0] + ix, ny = m[1] + iy;
                        v3 dp = dp0;
                        dp.x += nx * L[0];
                        dp.y += ny * L[1];
                        v3 f;
                        if (same && nx == 0 && ny == 0) {
                            /* the molecule itself: getNonBond's intramolecular rule */
                            if (ja == ia || ja == ng[0] || ja == ng[1] || ja == ng[2] || ja == ng[3]) continue;
                            e += orc_ljqh_gpu(dp, ri[0] + rj[0], ri[1] * rj[1], ri[2] * rj[2], r2damp, &f);
                        } else {
                            if (v3_dot(dp, dp) >= s->rcut_inter2) continue;
                            e += orc_inter_pair(s, dp, ri, rj, r2damp, &f);
                        }
                        fe = v3_add(fe, f);
                        ORC_SCALE(s, ia, f);
                    }
                }
            }
            etot += e;
            v3_st(fapos + 3 * ia, v3_add(v3_ld(fapos + 3 * ia), fe));
        }
        return etot;
    }
    if (s->nb_mode == ORC_NB_GPU) {
        for (int ia = 0; ia < na; ia++) {
            v3 pi = v3_ld(apos + 3 * ia);
            const double *ri = s->reqs + 4 * ia;
            const int *ng = s->neighs + 4 * ia;
            v3 fe = {0, 0, 0};
            double e = 0.0;
            for (int ja = 0; ja < na; ja++) {
                if (ja == ia) continue;
                if (ja == ng[0] || ja == ng[1] || ja == ng[2] || ja == ng[3]) continue;
                const double *rj = s->reqs + 4 * ja;
                v3 dp = v3_sub(v3_ld(apos + 3 * ja), pi);
                v3 f;
                if (s->mol_id && s->mol_id[ja] != s->mol_id[ia]) {
                    if (v3_dot(dp, dp) >= s->rcut_inter2) continue;
                    e += orc_inter_pair(s, dp, ri, rj, r2damp, &f);
                } else {
                    /* mixREQ_arithmetic: R0=Ri+Rj, E0=Ei*Ej (pre-sqrt), Q=Qi*Qj (UFF.cl:57-65) */
                    e += orc_ljqh_gpu(dp, ri[0] + rj[0], ri[1] * rj[1], ri[2] * rj[2], r2damp, &f);
                }
                fe = v3_add(fe, f);
                ORC_SCALE(s, ia, f);
            }
            etot += e;
            v3_st(fapos + 3 * ia, v3_add(v3_ld(fapos + 3 * ia), fe));
        }
    } else if (s->nb_mode == ORC_NB_CPU) {
        double f2max = s->fmax_nb * s->fmax_nb;
        for (int ia = 0; ia < na; ia++) {
            v3 pi = v3_ld(apos + 3 * ia);
            const double *reqi = s->reqs + 4 * ia;
            int i0_ex = ia * ORC_EXCL_MAX;
            int iex = i0_ex;
            int iex_end = i0_ex + ORC_EXCL_MAX - 1;
            int jex = s->excl[iex];
            v3 fi = {0, 0, 0};
            double e = 0.0;
            for (int ja = 0; ja < na; ja++) {
                if (ja == ia) continue;
                if (jex != -1) {
                    if (iex < iex_end && jex < ja) iex++;
                    jex = s->excl[iex];
                }
                if (jex == ja) continue;
                v3 dp = v3_sub(v3_ld(apos + 3 * ja), pi);
                if (v3_dot(dp, dp) > s->rcut2) continue;
                const double *reqj = s->reqs + 4 * ja;
                /* combine_req(reqj, reqi) (nonbonded.rs:166-168) */
                double rq[4] = {reqj[0] + reqi[0], reqj[1] * reqi[1], reqj[2] * reqi[2], reqj[3] * reqi[3]};
                v3 f;
                double eij = orc_ljqh_cpu(dp, rq, r2damp, &f);
                if (s->clamp_nb) {
                    double ff2 = v3_dot(f, f);
                    if (ff2 > f2max) f = v3_mul(f, sqrt(f2max) / sqrt(ff2));
                }
                e += eij;
                fi = v3_add(fi, f);
                ORC_SCALE(s, ia, f);
            }
            etot += e;
            v3_st(fapos + 3 * ia, v3_add(v3_ld(fapos + 3 * ia), fi));
        }
    }
    return etot;
}

/* NonBondedFF::make_second_neighs (molff/src/nonbonded.rs:112-137): sorted
 * 1-2 and 1-3 exclusion lists.  Returns 0, or -1-ia if an atom overflows.   */
int orc_make_second_neighs(int natoms, const int *neighs, int *excl) {
    for (int ia = 0; ia < natoms; ia++) {
        int *ex = excl + ia * ORC_EXCL_MAX;
        for (int k = 0; k < ORC_EXCL_MAX; k++) ex[k] = -1;
        int n = 0;
        int cand[20];
        int nc = 0;
        for (int s = 0; s < 4; s++) cand[nc++] = neighs[4 * ia + s];
        for (int s = 0; s < 4; s++) {
            int ja = neighs[4 * ia + s];
            if (ja < 0 || ja >= natoms) continue;
            for (int t = 0; t < 4; t++) cand[nc++] = neighs[4 * ja + t];
        }
        for (int c = 0; c < nc; c++) {
            int jb = cand[c];
            if (jb < 0 || jb == ia || jb >= natoms) continue;
            int dup = 0;
            for (int m = 0; m < n; m++) if (ex[m] == jb) { dup = 1; break; }
            if (dup) continue;
            if (n < ORC_EXCL_MAX) ex[n++] = jb;
            else return -1 - ia;
        }
        /* sort_unstable of distinct ints == any sort */
        for (int a = 1; a < n; a++) {
            int v = ex[a], b = a - 1;
            while (b >= 0 && ex[b] > v) { ex[b + 1] = ex[b]; b--; }
            ex[b + 1] = v;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* GridFF tricubic B-spline evaluation                                 */
/* ------------------------------------------------------------------ */

/* basis / dbasis  (opencl/gridff_eval.cl:48-69; UFF.cl:1416-1438) */
static inline void orc_basis(double u, double *b) {
    const double inv6 = 1.0 / 6.0;
    double u2 = u * u, t = 1.0 - u;
    b[0] = inv6 * t * t * t;
    b[1] = inv6 * (3.0 * u2 * (u - 2.0) + 4.0);
    b[2] = inv6 * (3.0 * u * (1.0 + u - u2) + 1.0);
    b[3] = inv6 * u2 * u;
}
static inline void orc_dbasis(double u, double *d) {
    double u2 = u * u, t = 1.0 - u;
    d[0] = -0.5 * t * t;
    d[1] = 0.5 * (3.0 * u2 - 4.0 * u);
    d[2] = 0.5 * (-3.0 * u2 + 2.0 * u + 1.0);
    d[3] = 0.5 * u2;
}
static inline int orc_modulo(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }

/* fe3d_pbc_comb (opencl/gridff_eval.cl:234-286; UFF.cl:1459-1511).  u is the
 * grid coordinate (p - p0)*invStep; returns (dE/du_x, dE/du_y, dE/du_z, E) of
 * E(u) = sum_stencil Bx By Bz (PLQH . C).  x,y wrap periodically; the result
 * is zero unless 1 <= iz < nz-2 (reference convention).                    */
void orc_fe3d_pbc_comb(const double *u, int nx, int ny, int nz, const double *C, const double *plqh, double *out) {
    int ix = (int)u[0], iy = (int)u[1], iz = (int)u[2];
    if (u[0] < 0) ix--;
    if (u[1] < 0) iy--;
    double tx = u[0] - ix, ty = u[1] - iy, tz = u[2] - iz;
    out[0] = out[1] = out[2] = out[3] = 0.0;
    if (iz < 1 || iz >= nz - 2) return;
    int x0 = orc_modulo(ix - 1, nx);
    int y0 = orc_modulo(iy - 1, ny);
    double bx[4], dx[4], by[4], dy[4], bz[4], dz[4];
    orc_basis(tx, bx); orc_dbasis(tx, dx);
    orc_basis(ty, by); orc_dbasis(ty, dy);
    orc_basis(tz, bz); orc_dbasis(tz, dz);
    /* same contraction order as the reference: z first (fe1Dcomb), then y
     * (fe2d_comb), then x (fe3d_pbc_comb).                                 */
    double Ex[4][3];
    for (int a = 0; a < 4; a++) {
        int xi = (x0 + a) % nx;
        double fe[4][2];
        for (int b = 0; b < 4; b++) {
            int yi = (y0 + b) % ny;
            const double *col = C + 4 * ((size_t)(iz - 1) + (size_t)nz * ((size_t)yi + (size_t)ny * (size_t)xi));
            double cs[4];
            for (int c = 0; c < 4; c++) {
                const double *E = col + 4 * c;
                cs[c] = plqh[0] * E[0] + plqh[1] * E[1] + plqh[2] * E[2] + plqh[3] * E[3];
            }
            fe[b][0] = bz[0] * cs[0] + bz[1] * cs[1] + bz[2] * cs[2] + bz[3] * cs[3];
            fe[b][1] = dz[0] * cs[0] + dz[1] * cs[1] + dz[2] * cs[2] + dz[3] * cs[3];
        }
        Ex[a][0] = fe[0][0] * dy[0] + fe[1][0] * dy[1] + fe[2][0] * dy[2] + fe[3][0] * dy[3];
        Ex[a][1] = fe[0][1] * by[0] + fe[1][1] * by[1] + fe[2][1] * by[2] + fe[3][1] * by[3];
        Ex[a][2] = fe[0][0] * by[0] + fe[1][0] * by[1] + fe[2][0] * by[2] + fe[3][0] * by[3];
    }
    out[0] = dx[0] * Ex[0][2] + dx[1] * Ex[1][2] + dx[2] * Ex[2][2] + dx[3] * Ex[3][2];
    out[1] = bx[0] * Ex[0][0] + bx[1] * Ex[1][0] + bx[2] * Ex[2][0] + bx[3] * Ex[3][0];
    out[2] = bx[0] * Ex[0][1] + bx[1] * Ex[1][1] + bx[2] * Ex[2][1] + bx[3] * Ex[3][1];
    out[3] = bx[0] * Ex[0][2] + bx[1] * Ex[1][2] + bx[2] * Ex[2][2] + bx[3] * Ex[3][2];
}

/* Per-atom PLQH mixing vector (opencl/UFF.cl:1681-1687):
 * ej = exp(alpha*R), PLQH = (ej^2*sqrtE, ej*sqrtE, Q, 0).                  */
void orc_plqh(const double *req, double alpha, double *plqh) {
    double ej = exp(alpha * req[0]);
    plqh[0] = ej * ej * req[1];
    plqh[1] = ej * req[1];
    plqh[2] = req[2];
    plqh[3] = 0.0;
}

/* Surface term of getNonBond_GridFF_Bspline (opencl/UFF.cl:1673-1711):
 * F = -invStep * dE/du.  Adds into fapos, returns sum of E.               */
double orc_grid_eval(const orc_system *s, const double *apos, double *fapos) {
    if (!s->grid) return 0.0;
    double etot = 0.0;
    for (int ia = 0; ia < s->natoms; ia++) {
        double plqh[4], u[3], fe[4];
        orc_plqh(s->reqs + 4 * ia, s->alpha_morse, plqh);
        for (int k = 0; k < 3; k++) u[k] = (apos[3 * ia + k] - s->gp0[k]) * s->ginv_step[k];
        orc_fe3d_pbc_comb(u, s->gnx, s->gny, s->gnz, s->grid, plqh, fe);
        v3 fs = {-fe[0] * s->ginv_step[0], -fe[1] * s->ginv_step[1], -fe[2] * s->ginv_step[2]};
        fapos[3 * ia + 0] += fs.x;
        fapos[3 * ia + 1] += fs.y;
        fapos[3 * ia + 2] += fs.z;
        ORC_SCALE(s, ia, fs);
        etot += fe[3];
    }
    return etot;
}

/* MolWorld::eval_forces (surfmol/src/mol_world.rs:60-86) with the GridFF
 * B-spline surface in place of the folded-basis surface.  E6 =
 * (eb, ea, ed, ei, enb, esurf).                                            */
void orc_eval_forces(const orc_system *s, const double *apos, double *fapos, double *E6) {
    double E4[4];
    orc_uff_eval(s, apos, fapos, E4);
    double enb = orc_nb_eval(s, apos, fapos);
    double es = orc_grid_eval(s, apos, fapos);
    E6[0] = E4[0]; E6[1] = E4[1]; E6[2] = E4[2]; E6[3] = E4[3]; E6[4] = enb; E6[5] = es;
}

/* ------------------------------------------------------------------ */
/* Integrators                                                          */
/* ------------------------------------------------------------------ */

/* One FIRE step on given forces f (modified in place by the clamp), the
 * body of the loop of surfmol/tests/relax_pentacene_uff.rs:142-193 with
 * FireState semantics from molff/src/raff.rs:1040-1062:
 *   v += F dt; sums of v.F, |v|^2 over the updated v; if v.F > 0 and
 *   |F|^2, |v|^2 > 1e-30: v = (1-a) v + a |v| F/|F|, n_pos++, and after
 *   n_min such steps dt = min(f_inc dt, dt_max), a *= f_alpha; else n_pos = 0,
 *   dt *= f_dec, a = a0, v = 0; finally x += v dt with the dt the step began
 *   with.
 * flim > 0 clamps each |F_i| to flim as step_fire does (raff.rs:1089-1096):
 * the clamped F enters the velocity update, v.F and the mixing direction,
 * while |F|^2 (and so |F| in the mixing) sums the unclamped forces.
 * state[3] = (dt, alpha, n_pos), updated.  Returns max_i |F_i|^2 of the
 * unclamped forces.                                                     */
double orc_fire_step(int na, double *f, double *vapos, double *apos, double *state, const orc_fire_params *fp) {
    double dt = state[0], alpha = state[1];
    int n_pos = (int)state[2];
    double v_dot_f = 0.0, v_norm2 = 0.0, f_norm2 = 0.0, f2max = 0.0;
    for (int ia = 0; ia < na; ia++) {
        v3 fi = v3_ld(f + 3 * ia);
        double f2 = v3_dot(fi, fi);
        if (f2 > f2max) f2max = f2;
        if (fp->flim > 0.0 && f2 > fp->flim * fp->flim) {
            fi = v3_mul(fi, fp->flim / sqrt(f2));
            v3_st(f + 3 * ia, fi);
        }
        f_norm2 += f2;
        v3 v = v3_ld(vapos + 3 * ia);
        v = v3_add(v, v3_mul(fi, dt));
        v3_st(vapos + 3 * ia, v);
        v_dot_f += v3_dot(v, fi);
        v_norm2 += v3_dot(v, v);
    }
    double dt_next = dt;
    if (v_dot_f > 0.0 && f_norm2 > 1e-30 && v_norm2 > 1e-30) {
        double v_mag = sqrt(v_norm2);
        double f_mag = sqrt(f_norm2);
        for (int ia = 0; ia < na; ia++) {
            v3 fh = v3_mul(v3_ld(f + 3 * ia), 1.0 / f_mag);
            v3 v = v3_ld(vapos + 3 * ia);
            v3_st(vapos + 3 * ia, v3_lincomb(1.0 - alpha, v, alpha * v_mag, fh));
        }
        n_pos++;
        if (n_pos > fp->n_min) {
            dt_next = fmin(dt * fp->f_inc, fp->dt_max);
            alpha *= fp->f_alpha;
        }
    } else {
        n_pos = 0;
        dt_next = dt * fp->f_dec;
        alpha = fp->alpha0;
        for (int i = 0; i < 3 * na; i++) vapos[i] = 0.0;
    }
    for (int ia = 0; ia < na; ia++) {
        v3 p = v3_ld(apos + 3 * ia);
        v3_st(apos + 3 * ia, v3_add(p, v3_mul(v3_ld(vapos + 3 * ia), dt)));
    }
    state[0] = dt_next;
    state[1] = alpha;
    state[2] = (double)n_pos;
    return f2max;
}

/* FIRE loop of surfmol/tests/relax_pentacene_uff.rs:142-193 (identical to
 * relax_xylitol_uff.rs:43-100 and molengine relax_fire main.rs:237-282):
 * evaluate the forces, take orc_fire_step, stop when max_i |F_i| of the
 * step's (unclamped) forces is below fconv -- tested after the move.
 *
 * Returns the number of steps taken (itr+1 at convergence, max_steps if not).
 * out[0] = total energy at the last force evaluation (sum of E6),
 * out[1] = max_i |F_i| at the last evaluation, out[2..7] = E6 terms,
 * out[8] = 1 if converged.  apos/vapos are updated in place (the last move is
 * applied, as in the reference loop).                                      */
int orc_relax_fire(const orc_system *s, double *apos, double *vapos, const orc_fire_params *fp,
                   int max_steps, double fconv, double *out) {
    int na = s->natoms;
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)na);
    double state[3] = {fp->dt0, fp->alpha0, 0.0};
    int steps = max_steps, conv = 0;
    double E6[6] = {0, 0, 0, 0, 0, 0}, fmax = 0.0;
    for (int itr = 0; itr < max_steps; itr++) {
        orc_eval_forces(s, apos, f, E6);
        fmax = sqrt(orc_fire_step(na, f, vapos, apos, state, fp));
        if (fmax < fconv) { steps = itr + 1; conv = 1; break; }
    }
    out[0] = E6[0] + E6[1] + E6[2] + E6[3] + E6[4] + E6[5];
    out[1] = fmax;
    for (int k = 0; k < 6; k++) out[2 + k] = E6[k];
    out[8] = conv;
    free(f);
    return steps;
}

/* One damped-MD step of MolWorld::run_md (surfmol/src/mol_world.rs:145-160)
 * on given forces: DynamicAtoms::move_atom_md (moltopo/src/molecular.rs:
 * 118-137) per atom -- the sums take the velocity before the update and the
 * unclamped force; |F| is clamped to flim for the update;
 * v = cdamp v + F_clamped dt; x += v dt -- then, after every atom has moved,
 * all velocities are zeroed when sum v.F < 0.  sums[3] = (sum v.F, sum |v|^2,
 * sum |F|^2).  (run_md names these ff, vv and vf: mol_world.rs:150-157.)   */
void orc_md_step(int na, const double *f, double *vapos, double *apos, double dt, double flim, double cdamp,
                 double *sums) {
    double vdotf = 0.0, vv = 0.0, ff = 0.0;
    for (int i = 0; i < na; i++) {
        v3 fi = v3_ld(f + 3 * i), v = v3_ld(vapos + 3 * i), p = v3_ld(apos + 3 * i);
        double a = v3_dot(v, fi), b = v3_dot(v, v), f2 = v3_dot(fi, fi);
        v3 fc = fi;
        if (f2 > flim * flim) fc = v3_mul(fc, flim / sqrt(f2));
        v3 vn = v3_add(v3_mul(v, cdamp), v3_mul(fc, dt));
        v3 pn = v3_add(p, v3_mul(vn, dt));
        v3_st(apos + 3 * i, pn);
        v3_st(vapos + 3 * i, vn);
        vdotf += a; vv += b; ff += f2;
    }
    if (vdotf < 0.0) for (int i = 0; i < 3 * na; i++) vapos[i] = 0.0;
    sums[0] = vdotf; sums[1] = vv; sums[2] = ff;
}

/* MolWorld::run_md (surfmol/src/mol_world.rs:141-164): damped MD, velocities
 * zeroed when sum v.F < 0, stop when sum |F|^2 < fconv^2 (tested after the
 * move).  cdamp = max(0, 1 - damping).  Returns steps taken.               */
int orc_run_md(const orc_system *s, double *apos, double *vapos, int niter, double dt, double fconv,
               double flim, double damping) {
    int na = s->natoms;
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)na);
    double f2conv = fconv * fconv;
    double cdamp = 1.0 - damping;
    if (cdamp < 0.0) cdamp = 0.0;
    int ret = niter;
    for (int itr = 0; itr < niter; itr++) {
        double E6[6], sums[3];
        orc_eval_forces(s, apos, f, E6);
        orc_md_step(na, f, vapos, apos, dt, flim, cdamp, sums);
        if (sums[2] < f2conv) { ret = itr + 1; break; }
    }
    free(f);
    return ret;
}

/* n plain MD steps of DynamicAtoms::move_atom_md (moltopo/src/molecular.rs:
 * 118-137) with no velocity purge and no convergence test, as driven by the
 * invariants loop of surfmol/tests/relax_xylitol_uff.rs:183-198.           */
void orc_md_steps(const orc_system *s, double *apos, double *vapos, int nsteps, double dt, double flim, double cdamp) {
    int na = s->natoms;
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)na);
    for (int it = 0; it < nsteps; it++) {
        double E6[6];
        orc_eval_forces(s, apos, f, E6);
        for (int i = 0; i < na; i++) {
            v3 fi = v3_ld(f + 3 * i), v = v3_ld(vapos + 3 * i), p = v3_ld(apos + 3 * i);
            double f2 = v3_dot(fi, fi);
            if (f2 > flim * flim) fi = v3_mul(fi, flim / sqrt(f2));
            v3 vn = v3_add(v3_mul(v, cdamp), v3_mul(fi, dt));
            v3_st(vapos + 3 * i, vn);
            v3_st(apos + 3 * i, v3_add(p, v3_mul(vn, dt)));
        }
    }
    free(f);
}

/* ------------------------------------------------------------------ */
/* GridFF reference-grid build and B-spline fit                         */
/* ------------------------------------------------------------------ */

/* make_GridFF (opencl/gridff_build.cl:908-1062) in FP64.  Node (ia,ib,ic) sits
 * at p0 + lvec_a*ia/nx + lvec_b*ib/ny + lvec_c*ic/nz; the substrate atoms are
 * summed over the fixed image set |ix|<=npbc[0], |iy|<=npbc[1], |iz|<=npbc[2]
 * of the cell (lvec).  Per substrate atom REQ = (R, sqrt(E), Q):
 *   Pauli  += sqrtE * exp(-2 alpha (r - R))
 *   London += -2 sqrtE * exp(-alpha (r - R))
 *   Coul   += kC * Q / sqrt(r^2 + rdamp^2)
 * Output: out[4*(iz + nz*(iy + ny*ix)) + {0,1,2}] (z-fastest, channel 3 = 0),
 * i.e. already in the packed layout of GridFF.h pack_Bspline_d.            */
/* Reference-grid values (Pauli, London, Coulomb) at node (ix, iy, iz): the
 * body of make_GridFF's per-node loop (opencl/gridff_build.cl:1006-1052). */
static void orc_gridff_node(int natoms, const double *apos, const double *reqs, const double *lvec,
                            int nx, int ny, int nz, const double *p0, const int *npbc, double alpha,
                            double r2damp, int ix, int iy, int iz, double *o) {
    double pos[3];
    for (int k = 0; k < 3; k++)
        pos[k] = p0[k] + lvec[0 + k] * ix / (double)nx + lvec[3 + k] * iy / (double)ny + lvec[6 + k] * iz / (double)nz;
    double eP = 0.0, eL = 0.0, eC = 0.0;
    for (int ja = 0; ja < natoms; ja++) {
        const double *R = reqs + 4 * ja;
        for (int jz = -npbc[2]; jz <= npbc[2]; jz++)
            for (int jy = -npbc[1]; jy <= npbc[1]; jy++)
                for (int jx = -npbc[0]; jx <= npbc[0]; jx++) {
                    double d[3];
                    for (int k = 0; k < 3; k++)
                        d[k] = pos[k] - apos[3 * ja + k] + lvec[0 + k] * jx + lvec[3 + k] * jy + lvec[6 + k] * jz;
                    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    double r = sqrt(r2 + 1e-32);
                    double ir2 = 1.0 / (r2 + r2damp);
                    eC += ORC_COULOMB_CONST * R[2] * sqrt(ir2);
                    double e = exp(-alpha * (r - R[0]));
                    double eM = R[1] * e;
                    eP += eM * e;
                    eL += -2.0 * eM;
                }
    }
    o[0] = eP; o[1] = eL; o[2] = eC; o[3] = 0.0;
}

void orc_gridff_build(int natoms, const double *apos, const double *reqs, const double *lvec,
                      int nx, int ny, int nz, const double *p0, const int *npbc,
                      double alpha, double rdamp, double *out) {
    double r2damp = rdamp * rdamp;
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
            for (int iz = 0; iz < nz; iz++)
                orc_gridff_node(natoms, apos, reqs, lvec, nx, ny, nz, p0, npbc, alpha, r2damp, ix, iy, iz,
                                out + 4 * ((size_t)iz + (size_t)nz * ((size_t)iy + (size_t)ny * (size_t)ix)));
}

/* The same values at a list of nodes only: nodes[3*q] = (ix, iy, iz),
 * out[4*q] = (Pauli, London, Coulomb, 0) -- spot checks of full-size grids. */
void orc_gridff_build_nodes(int natoms, const double *apos, const double *reqs, const double *lvec,
                            int nx, int ny, int nz, const double *p0, const int *npbc,
                            double alpha, double rdamp, int nnodes, const int *nodes, double *out) {
    double r2damp = rdamp * rdamp;
    for (int q = 0; q < nnodes; q++)
        orc_gridff_node(natoms, apos, reqs, lvec, nx, ny, nz, p0, npbc, alpha, r2damp, nodes[3 * q],
                        nodes[3 * q + 1], nodes[3 * q + 2], out + 4 * (size_t)q);
}

/* Apply the B-spline node operator S = Sx (x) Sy (x) Sz with the 1-D stencil
 * [1/6, 2/3, 1/6] (BsplineConv3D, opencl/gridff_build.cl:8-61 and
 * GridFF.cl:561-614): periodic in x and y, truncated (zero outside) in z.
 * Layout z-fastest with 4 channels; all 4 channels are processed.          */
void orc_bspline_apply(int nx, int ny, int nz, const double *C, double *out) {
    const double B0 = 2.0 / 3.0, B1 = 1.0 / 6.0;
    size_t n = (size_t)nx * ny * nz * 4;
    double *t1 = (double *)malloc(sizeof(double) * n);
    double *t2 = (double *)malloc(sizeof(double) * n);
#define IDX(ix, iy, iz) (4 * ((size_t)(iz) + (size_t)nz * ((size_t)(iy) + (size_t)ny * (size_t)(ix))))
    for (int ix = 0; ix < nx; ix++) for (int iy = 0; iy < ny; iy++) for (int iz = 0; iz < nz; iz++)
        for (int c = 0; c < 4; c++) {
            double v = B0 * C[IDX(ix, iy, iz) + c];
            if (iz > 0) v += B1 * C[IDX(ix, iy, iz - 1) + c];
            if (iz < nz - 1) v += B1 * C[IDX(ix, iy, iz + 1) + c];
            t1[IDX(ix, iy, iz) + c] = v;
        }
    for (int ix = 0; ix < nx; ix++) for (int iy = 0; iy < ny; iy++) for (int iz = 0; iz < nz; iz++)
        for (int c = 0; c < 4; c++)
            t2[IDX(ix, iy, iz) + c] = B0 * t1[IDX(ix, iy, iz) + c] + B1 * t1[IDX(ix, orc_modulo(iy - 1, ny), iz) + c] +
                                      B1 * t1[IDX(ix, (iy + 1) % ny, iz) + c];
    for (int ix = 0; ix < nx; ix++) for (int iy = 0; iy < ny; iy++) for (int iz = 0; iz < nz; iz++)
        for (int c = 0; c < 4; c++)
            out[IDX(ix, iy, iz) + c] = B0 * t2[IDX(ix, iy, iz) + c] + B1 * t2[IDX(orc_modulo(ix - 1, nx), iy, iz) + c] +
                                       B1 * t2[IDX((ix + 1) % nx, iy, iz) + c];
#undef IDX
    free(t1); free(t2);
}

/* Dense Gaussian elimination with partial pivoting on an n x n system (FP64).
 * Used by the oracle's fit as an independent, simple solver for the 1-D
 * B-spline interpolation matrices (the product uses Thomas / Sherman-Morrison
 * sweeps).  A is overwritten.                                              */
static void orc_dense_solve(int n, double *A, double *b) {
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++) if (fabs(A[i * n + k]) > fabs(A[p * n + k])) p = i;
        if (p != k) {
            for (int j = 0; j < n; j++) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
            double t = b[k]; b[k] = b[p]; b[p] = t;
        }
        for (int i = k + 1; i < n; i++) {
            double m = A[i * n + k] / A[k * n + k];
            if (m == 0.0) continue;
            for (int j = k; j < n; j++) A[i * n + j] -= m * A[k * n + j];
            b[i] -= m * b[k];
        }
    }
    for (int k = n - 1; k >= 0; k--) {
        double s = b[k];
        for (int j = k + 1; j < n; j++) s -= A[k * n + j] * b[j];
        b[k] = s / A[k * n + k];
    }
}

/* Exact FP64 B-spline fit: find C with S C = E for the operator of
 * orc_bspline_apply.  The reference's iterative driver (FireCore GridFF.h,
 * external to the reference repository) converges to this same solution;
 * here each 1-D factor is inverted directly by dense elimination (the
 * oracle favours obviously-correct over fast).  E and C are z-fastest with
 * 4 channels; C may alias E.                                               */
void orc_bspline_fit(int nx, int ny, int nz, const double *E, double *C) {
    const double B0 = 2.0 / 3.0, B1 = 1.0 / 6.0;
    size_t n = (size_t)nx * ny * nz * 4;
    if (C != E) memcpy(C, E, sizeof(double) * n);
#define IDX(ix, iy, iz) (4 * ((size_t)(iz) + (size_t)nz * ((size_t)(iy) + (size_t)ny * (size_t)(ix))))
    int nmax = nx > ny ? nx : ny; if (nz > nmax) nmax = nz;
    double *A = (double *)malloc(sizeof(double) * (size_t)nmax * nmax);
    double *b = (double *)malloc(sizeof(double) * (size_t)nmax);
    /* z lines: tridiagonal, truncated */
    for (int ix = 0; ix < nx; ix++) for (int iy = 0; iy < ny; iy++) for (int c = 0; c < 4; c++) {
        memset(A, 0, sizeof(double) * (size_t)nz * nz);
        for (int i = 0; i < nz; i++) {
            A[i * nz + i] = B0;
            if (i > 0) A[i * nz + i - 1] = B1;
            if (i < nz - 1) A[i * nz + i + 1] = B1;
            b[i] = C[IDX(ix, iy, i) + c];
        }
        orc_dense_solve(nz, A, b);
        for (int i = 0; i < nz; i++) C[IDX(ix, iy, i) + c] = b[i];
    }
    /* y lines: cyclic */
    for (int ix = 0; ix < nx; ix++) for (int iz = 0; iz < nz; iz++) for (int c = 0; c < 4; c++) {
        memset(A, 0, sizeof(double) * (size_t)ny * ny);
        for (int i = 0; i < ny; i++) {
            A[i * ny + i] += B0;
            A[i * ny + orc_modulo(i - 1, ny)] += B1;
            A[i * ny + (i + 1) % ny] += B1;
            b[i] = C[IDX(ix, i, iz) + c];
        }
        orc_dense_solve(ny, A, b);
        for (int i = 0; i < ny; i++) C[IDX(ix, i, iz) + c] = b[i];
    }
    /* x lines: cyclic */
    for (int iy = 0; iy < ny; iy++) for (int iz = 0; iz < nz; iz++) for (int c = 0; c < 4; c++) {
        memset(A, 0, sizeof(double) * (size_t)nx * nx);
        for (int i = 0; i < nx; i++) {
            A[i * nx + i] += B0;
            A[i * nx + orc_modulo(i - 1, nx)] += B1;
            A[i * nx + (i + 1) % nx] += B1;
            b[i] = C[IDX(i, iy, iz) + c];
        }
        orc_dense_solve(nx, A, b);
        for (int i = 0; i < nx; i++) C[IDX(i, iy, iz) + c] = b[i];
    }
#undef IDX
    free(A); free(b);
}

/* ------------------------------------------------------------------ */
/* Multi-replica rigid-atom port relaxation (RAFF "RRsp3" cluster step) */
/* ------------------------------------------------------------------ */

/* Shared topology of one replica (oclff::RRsp3Multi uploads,
 * crates/libs/oclff/src/rrsp3.rs:829-929). */
typedef struct {
    int n, nnode;
    const double *radius;      /* [n]                                   */
    const int *neighs;         /* [4n] local, -1 padded                 */
    const int *excl1, *excl2;  /* [4n]                                  */
    const int *bk_slots;       /* [4n] node-port index inode*4+k        */
    const int *fixmask;        /* [n] or NULL                           */
    const double *port_local;  /* [nnode*4*3] port arms (body frame)    */
    const double *kflat;       /* [nnode*4]                             */
    const double *inv_mass;    /* [n]; <= 1e-12: padding, never moves   */
} orc_raff;

static inline v3 orc_qrot(const double *q, v3 v) {
    v3 u = {q[0], q[1], q[2]};
    v3 t = v3_mul(v3_cross(u, v), 2.0);
    return v3_add(v3_add(v, v3_mul(t, q[3])), v3_cross(u, t));
}

static inline void orc_qmul(const double *a, const double *b, double *o) {
    double r[4] = {a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1],
                   a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0],
                   a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3],
                   a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2]};
    o[0] = r[0]; o[1] = r[1]; o[2] = r[2]; o[3] = r[3];
}

static inline void orc_qnormalize(double *q) {
    double n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (n2 < 1e-16) { q[0] = q[1] = q[2] = 0.0; q[3] = 1.0; return; }
    double s = 1.0 / sqrt(n2);
    q[0] *= s; q[1] *= s; q[2] *= s; q[3] *= s;
}

/* One step of RRsp3Multi::step_cluster_multi (rrsp3.rs:1000-1026) for one
 * replica, in FP64: compute_collision_multi (opencl/RRsp3.cl:1774-1831),
 * compute_ports_current_multi (:1835-1933) and apply_corrections_multi
 * (:1937-2047, massless_rot = 0), with the kernels' epsilons.  All reads of
 * the first two phases see the positions and orientations at the start of
 * the step.  Updates pos [3n], quat [4n] (x,y,z,w), mom [3n], qmom [4n];
 * writes this step's dcoll [3n], dnode / drot [3*nnode], dneigh
 * [12*nnode].  Returns the largest |component| of dcoll over the atoms and
 * of dnode over the nodes. */
double orc_raff_step(const orc_raff *t, double dt, double rot_mass_scale, double relaxation, double beta,
                     int collide, double *pos, double *quat, double *mom, double *qmom, double *dcoll,
                     double *dnode, double *drot, double *dneigh) {
    const int n = t->n, nn = t->nnode;
    double metric = 0.0;
    /* collision: Jacobi sphere-pair corrections */
    for (int i = 0; i < n; i++) {
        v3 sum = {0, 0, 0};
        const double wi = t->inv_mass[i], ri = t->radius[i];
        if (collide && wi > 1e-12 && ri > 0.0) {
            v3 p = v3_ld(pos + 3 * i);
            for (int j = 0; j < n; j++) {
                if (j == i) continue;
                int ex = 0;
                for (int k = 0; k < 4; k++) ex |= (t->excl1[4 * i + k] == j) | (t->excl2[4 * i + k] == j);
                if (ex) continue;
                const double wj = t->inv_mass[j], rj = t->radius[j];
                if (wj <= 1e-12 || rj <= 0.0) continue;
                v3 d = v3_sub(p, v3_ld(pos + 3 * j));
                double d2 = v3_dot(d, d), rs = ri + rj;
                if (d2 < rs * rs && d2 > 1e-16) {
                    double dist = sqrt(d2);
                    v3 nv = v3_mul(d, 1.0 / dist);
                    double dl = 0.5 * ((rs - dist) / (wi + wj + 1e-12));
                    sum = v3_add(sum, v3_mul(nv, dl * wi));
                }
            }
        }
        v3_st(dcoll + 3 * i, sum);
        double m = fmax(fabs(sum.x), fmax(fabs(sum.y), fabs(sum.z)));
        if (m > metric) metric = m;
    }
    /* ports (massfull XPBD, the Current variant) */
    const double dt2 = dt * dt + 1e-16;
    for (int i = 0; i < nn; i++) {
        v3 sdp = {0, 0, 0}, sdt = {0, 0, 0};
        const double wi = t->inv_mass[i];
        for (int k = 0; k < 4; k++) {
            v3 rec = {0, 0, 0};
            const int j = t->neighs[4 * i + k];
            if (wi > 1e-12 && j >= 0 && j < n && t->kflat[4 * i + k] > 0.0 && t->inv_mass[j] > 1e-12) {
                const double wj = t->inv_mass[j], K = t->kflat[4 * i + k];
                const double mi = 1.0 / wi, invI = rot_mass_scale / (0.4 * mi + 1e-12);
                v3 arm = {t->port_local[12 * i + 3 * k], t->port_local[12 * i + 3 * k + 1], t->port_local[12 * i + 3 * k + 2]};
                v3 ra = orc_qrot(quat + 4 * i, arm);
                v3 diff = v3_sub(v3_ld(pos + 3 * j), v3_add(v3_ld(pos + 3 * i), ra));
                double d2 = v3_dot(diff, diff);
                if (d2 >= 1e-16) {
                    double dist = sqrt(d2);
                    v3 nv = v3_mul(diff, 1.0 / dist);
                    v3 rxn = v3_cross(ra, nv);
                    double wang = v3_dot(rxn, rxn) * invI;
                    double alpha = 1.0 / (K * dt2);
                    int jnode = j < nn;
                    double wt = wi + wj + wang + alpha + 1e-12;
                    if (jnode) wt += wang;
                    double imp = dist / wt;
                    if (jnode) imp *= 0.5;
                    v3 P = v3_mul(nv, imp);
                    sdp = v3_add(sdp, v3_mul(P, wi));
                    sdt = v3_add(sdt, v3_mul(v3_cross(ra, P), invI));
                    rec = v3_mul(P, -wj);
                }
            }
            v3_st(dneigh + 12 * i + 3 * k, rec);
        }
        v3_st(dnode + 3 * i, sdp);
        v3_st(drot + 3 * i, sdt);
        double m = fmax(fabs(sdp.x), fmax(fabs(sdp.y), fabs(sdp.z)));
        if (m > metric) metric = m;
    }
    /* corrections */
    for (int i = 0; i < n; i++) {
        if (t->inv_mass[i] <= 1e-12) continue;
        const int isnode = i < nn;
        v3 dxp = isnode ? v3_ld(dnode + 3 * i) : (v3){0, 0, 0};
        for (int k = 0; k < 4; k++) {
            const int sl = t->bk_slots[4 * i + k];
            if (sl >= 0) dxp = v3_add(dxp, v3_ld(dneigh + 3 * sl));
        }
        v3 dx = v3_add(v3_ld(dcoll + 3 * i), dxp);
        const int msk = t->fixmask ? t->fixmask[i] : 0;
        if (msk & 1) dx.x = 0.0;
        if (msk & 2) dx.y = 0.0;
        if (msk & 4) dx.z = 0.0;
        v3 mv = v3_mul(dx, relaxation);
        if (beta != 0.0) mv = v3_add(mv, v3_mul(v3_ld(mom + 3 * i), beta));
        if (msk & 8) {
            double zold = pos[3 * i + 2];
            if (fabs(zold) > 1e-6) mv.z -= zold;
            else mv.z = 0.0;
        }
        v3_st(pos + 3 * i, v3_add(v3_ld(pos + 3 * i), mv));
        v3_st(mom + 3 * i, mv);
        if (!isnode || (msk & 7)) continue;
        v3 dth = v3_mul(v3_ld(drot + 3 * i), relaxation);
        double ang = sqrt(v3_dot(dth, dth));
        double *q = quat + 4 * i;
        double qold[4] = {q[0], q[1], q[2], q[3]}, qn[4] = {q[0], q[1], q[2], q[3]};
        if (ang > 1e-8) {
            /* quat_from_axis_angle with the axis dtheta/|dtheta| */
            v3 ax = v3_mul(dth, 1.0 / ang);
            double s = sin(0.5 * ang), c = cos(0.5 * ang);
            double dq[4] = {ax.x * s, ax.y * s, ax.z * s, c};
            orc_qmul(dq, qold, qn);
            orc_qnormalize(qn);
        }
        if (beta != 0.0)
            for (int k = 0; k < 4; k++) qn[k] += qmom[4 * i + k] * beta;
        orc_qnormalize(qn);
        if (qn[0] * qold[0] + qn[1] * qold[1] + qn[2] * qold[2] + qn[3] * qold[3] < 0.0)
            for (int k = 0; k < 4; k++) qn[k] = -qn[k];
        for (int k = 0; k < 4; k++) {
            qmom[4 * i + k] = qn[k] - qold[k];
            q[k] = qn[k];
        }
    }
    return metric;
}
