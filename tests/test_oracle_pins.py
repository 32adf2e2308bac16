"""Pins of the FP64 oracle (CPU): every function of oracle/ is checked against
a known answer from the reference's own tests/docs, a closed form, an
invariant, or a brute-force restatement in a different formulation -- chosen
so that a dropped term, a wrong sign or a swapped index fails."""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _two_atom_system(oracle, k=100.0, l0=0.9, r=1.0):
    O = oracle
    return O.OracleSystem(
        elems=["H", "H"], apos=np.array([[0.0, 0, 0], [r, 0, 0]]), bonds=np.array([[0, 1]], dtype=np.int32),
        angles=np.zeros((0, 3), np.int32), dihedrals=np.zeros((0, 4), np.int32), inversions=np.zeros((0, 4), np.int32),
        neighs=np.array([[1, -1, -1, -1], [0, -1, -1, -1]], np.int32),
        neigh_bs=np.array([[0, -1, -1, -1], [0, -1, -1, -1]], np.int32), types=["H_", "H_"],
        bon_params=np.array([[k, l0]]), ang_params=np.zeros((0, 5)), dih_params=np.zeros((0, 3)),
        inv_params=np.zeros((0, 4)), ang_ngs=np.zeros((0, 2), np.int32), dih_ngs=np.zeros((0, 3), np.int32),
        inv_ngs=np.zeros((0, 3), np.int32), reqs=np.zeros((2, 4)), nb_mode=O.NB_OFF)


def test_h2_bond_known_answer(oracle):
    """oclff/tests/test_uff_cl.rs:8-56: k=100, l0=0.9, r=1 -> Fx=20, E_atom=1."""
    s = _two_atom_system(oracle)
    f, E = oracle.uff_eval(s)
    assert f[0, 0] == pytest.approx(20.0, abs=1e-12)
    assert f[1, 0] == pytest.approx(-20.0, abs=1e-12)
    assert E[0] == pytest.approx(2.0, abs=1e-12)     # E_atom = 1 at each end
    assert np.abs(f[:, 1:]).max() == 0.0


def test_uff_bond_length_c_r(params, oracle):
    """C_R-C_R l0 = 1.458 A with bond order 1 (doc/topical_audit/uff.md:67)."""
    ti = params.atom_types["C_R"]
    ei = params.element_of_atom_type("C_R")
    l0 = oracle.uff_bond_length(ti, ti, ei, ei, oracle.bond_order_from_types("C_R", "C_R"))
    assert l0 == pytest.approx(1.458, abs=1e-12)


@pytest.fixture(scope="module")
def pentacene(oracle, params):
    el, x, q = oracle.read_xyz(os.path.join(GOLDEN, "pentacene.xyz"))
    return oracle.system_from_geometry(el, x, params, nb_mode=oracle.NB_OFF)


def test_pentacene_golden_energies(oracle, pentacene):
    """doc/topical_audit/uff.md:91: planar pentacene E_bond=3.83,
    E_angle=0.01, E_dih=0, E_inv=0; topology counts of the relax test."""
    s = pentacene
    assert (len(s.bonds), len(s.angles), len(s.dihedrals), len(s.inversions)) == (40, 66, 104, 66)
    f, E = oracle.uff_eval(s)
    assert E[0] == pytest.approx(3.83, abs=5e-3)
    assert E[1] == pytest.approx(0.01, abs=5e-4)
    assert abs(E[2]) < 1e-12 and abs(E[3]) < 1e-12


def test_pentacene_out_of_plane_force(oracle, pentacene):
    """relax_pentacene_uff.rs:97-106: atom 0 moved to z=0.5 feels a restoring
    force.  With the current topology code (each dihedral once) it is
    Fz = -8.3376 eV/A; counting every dihedral twice -- the state before the
    dedup fix listed in doc/topical_audit/uff.md:76 -- gives the documented
    -11.7 (uff.md:92)."""
    s = pentacene
    x = s.apos.copy()
    x[0, 2] = 0.5
    f, _ = oracle.uff_eval(s, x)
    assert f[0, 2] == pytest.approx(-8.3376, abs=5e-4)
    s2 = oracle.OracleSystem(**{**s.__dict__, "_keep": []})
    s2.dihedrals = np.concatenate([s.dihedrals, s.dihedrals])
    s2.dih_ngs = np.concatenate([s.dih_ngs, s.dih_ngs])
    s2.dih_params = np.concatenate([s.dih_params, s.dih_params])
    f2, _ = oracle.uff_eval(s2, x)
    assert f2[0, 2] == pytest.approx(-11.7, abs=0.05)


def _fd_grad(fun, x, h=1e-5):
    g = np.zeros_like(x)
    for i in range(x.shape[0]):
        for k in range(3):
            xp = x.copy(); xp[i, k] += h
            xm = x.copy(); xm[i, k] -= h
            g[i, k] = (fun(xp) - fun(xm)) / (2 * h)
    return g


def _angle_c1_term(s, x):
    """k*c1*cos(theta) summed over angles -- the term the reference's angle
    energy leaves out while its force keeps it (uff.rs:482-493)."""
    e = 0.0
    for (i, j, k), p in zip(s.angles, s.ang_params):
        a = x[i] - x[j]
        b = x[k] - x[j]
        c = a.dot(b) / np.linalg.norm(a) / np.linalg.norm(b)
        e += p[0] * p[2] * c
    return e


@pytest.mark.parametrize("name", ["xylitol", "PTCDA"])
def test_uff_force_is_minus_gradient(oracle, params, name):
    """F = -dE/dx for bonds+angles+dihedrals+inversions on a distorted
    molecule with real UFF params (sp3 centres in xylitol exercise c1!=0)."""
    el, x, q = oracle.read_xyz(os.path.join(GOLDEN, f"{name}.xyz"))
    s = oracle.system_from_geometry(el, x, params, nb_mode=oracle.NB_OFF)
    rng = np.random.default_rng(1)
    xd = x + rng.normal(0, 0.08, x.shape)

    def energy(xx):
        _, E = oracle.uff_eval(s, xx)
        # bond energy is counted at both ends: the physical bond energy is E/2
        return 0.5 * E[0] + E[1] + E[2] + E[3] + _angle_c1_term(s, xx)

    f, _ = oracle.uff_eval(s, xd)
    g = _fd_grad(energy, xd)
    err = np.abs(f + g).max() / np.abs(f).max()
    assert err < 1e-6, err
    if name == "xylitol":
        assert np.abs(s.ang_params[:, 2]).max() > 0.1   # sp3 angles have c1 != 0


def test_uff_forces_sum_to_zero(oracle, params):
    el, x, q = oracle.read_xyz(os.path.join(GOLDEN, "xylitol.xyz"))
    s = oracle.system_from_geometry(el, x, params, nb_mode=oracle.NB_OFF)
    xd = x + np.random.default_rng(2).normal(0, 0.1, x.shape)
    f, _ = oracle.uff_eval(s, xd)
    assert np.abs(f.sum(axis=0)).max() < 1e-10
    assert np.abs(np.cross(xd, f).sum(axis=0)).max() < 1e-9


def _pair(oracle, R, sE, Q, r, mode=None, rdamp=0.0):
    O = oracle
    s = _two_atom_system(oracle)
    s.neighs = -np.ones((2, 4), np.int32)
    s.neigh_bs = -np.ones((2, 4), np.int32)
    s.bonds = np.zeros((0, 2), np.int32)
    s.bon_params = np.zeros((0, 2))
    s.reqs = np.array([[R[0], sE[0], Q[0], 0.0], [R[1], sE[1], Q[1], 0.0]])
    s.nb_mode = O.NB_GPU if mode is None else mode
    s.rdamp = rdamp
    s.apos = np.array([[0.0, 0, 0], [r, 0, 0]])
    return s


def test_lj_closed_form_at_r0(oracle):
    """getLJQH at r = R0 = Ri+Rj: E = -E0 with E0 = sqrtEi*sqrtEj, F = 0."""
    s = _pair(oracle, (1.9, 1.5), (0.07, 0.04), (0.0, 0.0), 3.4)
    f, e = oracle.nb_eval(s)
    assert e == pytest.approx(-2 * 0.07 * 0.04, rel=1e-13)     # both ends
    assert np.abs(f).max() < 1e-14


def test_coulomb_closed_form(oracle):
    """Pure damped Coulomb: E_pair = kC qi qj / sqrt(r^2 + Rd^2)."""
    r, rd, qi, qj = 2.7, 1.0, 0.4, -0.3
    s = _pair(oracle, (1.0, 1.0), (0.0, 0.0), (qi, qj), r, rdamp=rd)
    f, e = oracle.nb_eval(s)
    ref = 14.3996448915 * qi * qj / math.sqrt(r * r + rd * rd)
    assert e == pytest.approx(2 * ref, rel=1e-13)
    fx = 14.3996448915 * qi * qj * r / (r * r + rd * rd) ** 1.5   # -dE/dr pushes atom 0 along -x for like charges
    assert f[0, 0] == pytest.approx(-fx, rel=1e-12)


def test_nb_force_is_minus_gradient(oracle, params):
    el, x, q = oracle.read_xyz(os.path.join(GOLDEN, "PTCDA.xyz"))
    rng = np.random.default_rng(3)
    qq = rng.normal(0, 0.2, len(el)); qq -= qq.mean()
    s = oracle.system_from_geometry(el, x, params, charges=qq, nb_mode=oracle.NB_GPU, rdamp=1.0)
    xd = x + rng.normal(0, 0.05, x.shape)
    f, _ = oracle.nb_eval(s, xd)
    g = _fd_grad(lambda xx: 0.5 * oracle.nb_eval(s, xx)[1], xd)
    assert np.abs(f + g).max() / np.abs(f).max() < 1e-6


def test_nb_conventions_agree_when_made_equal(oracle, params):
    """The CPU convention (NonBondedFF: E0=(a_i a_j)^2 with a=reqs[1], sorted
    1-2+1-3 exclusion walk, H term, clamp) reduces to the GPU convention
    (getNonBond: E0=a_i a_j, 1-2 only) when a -> sqrt(a), H=0, no clamp, no
    cutoff and the exclusion list holds 1-2 neighbours only."""
    el, x, q = oracle.read_xyz(os.path.join(GOLDEN, "PTCDA.xyz"))
    rng = np.random.default_rng(4)
    qq = rng.normal(0, 0.2, len(el)); qq -= qq.mean()
    sg = oracle.system_from_geometry(el, x, params, charges=qq, nb_mode=oracle.NB_GPU, rdamp=1.0)
    sc = oracle.system_from_geometry(el, x, params, charges=qq, nb_mode=oracle.NB_CPU, rdamp=1.0, clamp_nb=0)
    sc.reqs = sc.reqs.copy()
    sc.reqs[:, 1] = np.sqrt(sc.reqs[:, 1])
    sc.reqs[:, 3] = 0.0
    ex = np.full((len(el), 16), -1, np.int32)
    for i in range(len(el)):
        nn = sorted(int(j) for j in sc.neighs[i] if j >= 0)
        ex[i, :len(nn)] = nn
    sc.excl = ex
    fg, eg = oracle.nb_eval(sg)
    fc, ec = oracle.nb_eval(sc)
    assert ec == pytest.approx(eg, rel=1e-13)
    assert np.abs(fc - fg).max() < 1e-12
    # the full CPU convention excludes 1-3 pairs too: energies must differ
    sc2 = oracle.system_from_geometry(el, x, params, charges=qq, nb_mode=oracle.NB_CPU, rdamp=1.0)
    assert abs(oracle.nb_eval(sc2)[1] - eg) > 1.0


def test_second_neighs_counts(oracle, params):
    el, x, q = oracle.read_xyz(os.path.join(GOLDEN, "PTCDA.xyz"))
    s = oracle.system_from_geometry(el, x, params)
    ex = oracle.make_second_neighs(s.neighs)
    n = len(el)
    npairs = sum(1 for i in range(n) for j in range(i + 1, n)) - int((ex >= 0).sum()) // 2
    assert npairs == 585   # unique pairs left after 1-2 and 1-3 exclusion (PTCDA)
    assert n * (n - 1) // 2 - len(s.bonds) == 659


# ---------------------------------------------------------------- GridFF sampler
def _rand_grid(rng, n=(8, 9, 10)):
    return rng.normal(0, 1, n + (4,))


def test_bspline_partition_of_unity(oracle):
    ns = (7, 8, 9)
    C = np.zeros(ns + (4,)); C[..., 0] = 1.0
    rng = np.random.default_rng(5)
    for _ in range(20):
        u = np.array([rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(1.0, ns[2] - 2.0 - 1e-9)])
        out = oracle.fe3d_pbc_comb(u, C, ns, [1, 0, 0, 0])
        assert out[3] == pytest.approx(1.0, abs=1e-13)
        assert np.abs(out[:3]).max() < 1e-13


def test_bspline_reproduces_polynomials(oracle):
    """Cubic B-splines with coefficients C_k = k, k^2, k^3 reproduce u,
    u^2 + 1/3, u^3 + u (z axis, and x away from the periodic seam)."""
    ns = (12, 12, 14)
    rng = np.random.default_rng(6)
    k = np.arange(ns[2], dtype=float)
    for cz, fz, dfz in [(k, lambda u: u, lambda u: 1.0), (k ** 2, lambda u: u * u + 1 / 3, lambda u: 2 * u),
                        (k ** 3, lambda u: u ** 3 + u, lambda u: 3 * u * u + 1)]:
        C = np.zeros(ns + (4,)); C[..., 1] = cz[None, None, :]
        for _ in range(10):
            u = np.array([rng.uniform(0, 12), rng.uniform(0, 12), rng.uniform(1.0, ns[2] - 2.0 - 1e-9)])
            out = oracle.fe3d_pbc_comb(u, C, ns, [0, 1, 0, 0])
            assert out[3] == pytest.approx(fz(u[2]), rel=1e-12, abs=1e-12)
            assert out[2] == pytest.approx(dfz(u[2]), rel=1e-12, abs=1e-12)
    kx = np.arange(ns[0], dtype=float)
    C = np.zeros(ns + (4,)); C[..., 2] = kx[:, None, None]
    for _ in range(10):
        u = np.array([rng.uniform(2.0, ns[0] - 3.0), rng.uniform(0, 12), rng.uniform(1.0, 10.0)])
        out = oracle.fe3d_pbc_comb(u, C, ns, [0, 0, 1, 0])
        assert out[3] == pytest.approx(u[0], rel=1e-12)
        assert out[0] == pytest.approx(1.0, rel=1e-12)


def test_bspline_periodic_and_zero_outside(oracle):
    rng = np.random.default_rng(7)
    ns = (8, 9, 10)
    C = _rand_grid(rng, ns)
    plq = rng.normal(0, 1, 4)
    for _ in range(10):
        u = np.array([rng.uniform(0, 8), rng.uniform(0, 9), rng.uniform(1, 7.9)])
        a = oracle.fe3d_pbc_comb(u, C, ns, plq)
        b = oracle.fe3d_pbc_comb(u + np.array([3 * ns[0], -2 * ns[1], 0.0]), C, ns, plq)
        assert np.abs(a - b).max() < 1e-10
    assert np.abs(oracle.fe3d_pbc_comb(np.array([1.0, 1.0, 0.5]), C, ns, plq)).max() == 0.0
    assert np.abs(oracle.fe3d_pbc_comb(np.array([1.0, 1.0, ns[2] - 2.0]), C, ns, plq)).max() == 0.0


def test_bspline_gradient_fd(oracle):
    rng = np.random.default_rng(8)
    ns = (8, 9, 10)
    C = _rand_grid(rng, ns)
    plq = rng.normal(0, 1, 4)
    h = 1e-6
    for _ in range(10):
        u = np.array([rng.uniform(0, 8), rng.uniform(0, 9), rng.uniform(1.1, 7.8)])
        out = oracle.fe3d_pbc_comb(u, C, ns, plq)
        for k in range(3):
            up = u.copy(); up[k] += h
            um = u.copy(); um[k] -= h
            fd = (oracle.fe3d_pbc_comb(up, C, ns, plq)[3] - oracle.fe3d_pbc_comb(um, C, ns, plq)[3]) / (2 * h)
            assert out[k] == pytest.approx(fd, rel=1e-6, abs=1e-8)


def test_gridff_build_is_factorised_morse(oracle, params):
    """PLQH . (Pauli, London, Coulomb) of the built grid equals the direct
    pair sum of the mixed Morse potential E0[e^{-2a(r-R0)} - 2e^{-a(r-R0)}]
    (R0 = Ri+Rj, E0 = sqrtEi sqrtEj) plus damped Coulomb, over the image set."""
    rng = np.random.default_rng(9)
    L, Lz, ns = 6.0, 4.0, (5, 6, 4)
    sub = np.array([[0.5, 0.7, 0.0], [3.4, 2.9, -0.4]])
    sreq = np.array([[1.49, 0.036, 0.9, 0], [1.97, 0.099, -0.9, 0]])
    alpha, rd, npbc = 1.5, 1.0, (1, 1, 0)
    E = oracle.gridff_build(sub, sreq, [L, 0, 0, 0, L, 0, 0, 0, Lz], ns, (0.0, 0.0, 2.0), npbc, alpha, rd)
    probe = np.array([1.93, 0.067, -0.3, 0.0])
    plq = oracle.plqh(probe, alpha)
    for (ix, iy, iz) in [(0, 0, 0), (2, 3, 1), (4, 5, 3)]:
        p = np.array([ix * L / ns[0], iy * L / ns[1], 2.0 + iz * Lz / ns[2]])
        ref = 0.0
        for a, r in zip(sub, sreq):
            for jx in (-1, 0, 1):
                for jy in (-1, 0, 1):
                    d = p - a + np.array([jx * L, jy * L, 0.0])
                    rr = np.linalg.norm(d)
                    R0, E0 = probe[0] + r[0], probe[1] * r[1]
                    ref += E0 * (math.exp(-2 * alpha * (rr - R0)) - 2 * math.exp(-alpha * (rr - R0)))
                    ref += 14.3996448915 * probe[2] * r[2] / math.sqrt(rr * rr + rd * rd)
        got = plq[:3].dot(E[ix, iy, iz, :3])
        assert got == pytest.approx(ref, rel=1e-11)


def test_bspline_fit_inverts_stencil(oracle):
    rng = np.random.default_rng(10)
    ns = (6, 7, 8)
    C = _rand_grid(rng, ns)
    E = oracle.bspline_apply(C, ns)
    C2 = oracle.bspline_fit(E, ns)
    assert np.abs(C2 - C).max() < 1e-11
    # the spline through the fitted coefficients takes the node values
    u = np.array([2.0, 3.0, 4.0])
    out = oracle.fe3d_pbc_comb(u, C2, ns, [1, 0, 0, 0])
    assert out[3] == pytest.approx(E[2, 3, 4, 0], rel=1e-10)


# ---------------------------------------------------------------- integrators
def _xylitol_dummy(oracle, params):
    """relax_xylitol_uff.rs:110-128: real topology, set_dummy_params
    (bonds k=100, l0 = 0.95 l; angles/dihedrals/inversions off)."""
    el, x, q = oracle.read_xyz(os.path.join(GOLDEN, "xylitol.xyz"))
    s = oracle.system_from_geometry(el, x, params, nb_mode=oracle.NB_OFF)
    l = np.linalg.norm(x[s.bonds[:, 1]] - x[s.bonds[:, 0]], axis=1)
    s.bon_params = np.stack([np.full(len(l), 100.0), 0.95 * l], axis=1)
    s.ang_params = np.tile([0.0, 1.0, -1.0, 0.0, 0.0], (len(s.angles), 1))
    s.dih_params = np.tile([0.0, 1.0, 3.0], (len(s.dihedrals), 1))
    s.inv_params = np.tile([0.0, 1.0, -1.0, 0.0], (len(s.inversions), 1))
    return s, x


def _xorshift_stream(seed, n):
    """relax_xylitol_uff.rs:18-23."""
    out = []
    m = (1 << 64) - 1
    for _ in range(n):
        seed ^= (seed << 13) & m
        seed ^= seed >> 7
        seed ^= (seed << 17) & m
        out.append(((seed & 0x7fffffffffffff) * (2.0 / (1 << 52))) - 1.0)
    return seed, out


def test_md_conserves_momentum(oracle, params):
    """relax_xylitol_uff.rs:170-229: undamped MD conserves |dP|<1e-8,
    |dL|<1e-6 over 500 steps of dt=0.01."""
    s, x = _xylitol_dummy(oracle, params)
    seed, r = _xorshift_stream(0x12345678, 3 * len(x) * 2)
    xd = x + 0.3 * np.array(r[:3 * len(x)]).reshape(-1, 3)
    v = 0.1 * np.array(r[3 * len(x):]).reshape(-1, 3)
    v -= v.mean(axis=0)
    P0, L0 = v.sum(axis=0), np.cross(xd, v).sum(axis=0)
    x1, v1 = oracle.md_steps(s, 500, 0.01, 1e18, 1.0, xd, v)
    P1, L1 = v1.sum(axis=0), np.cross(x1, v1).sum(axis=0)
    assert np.linalg.norm(P1 - P0) < 1e-8
    assert np.linalg.norm(L1 - L0) < 1e-6


def test_fire_converges_same_energy_from_two_starts(oracle, params):
    """relax_xylitol_uff.rs:230-233: FIRE reaches fmax<1e-3, and two starts
    reach the same energy within 1e-6."""
    s, x = _xylitol_dummy(oracle, params)
    fp = oracle.FireParams(dt0=0.001, dt_max=0.02)
    seed, r = _xorshift_stream(0x12345678, 3 * len(x))
    x1 = x + 0.3 * np.array(r).reshape(-1, 3)
    seed, r2 = _xorshift_stream(0xABCDEF, 3 * len(x))
    x2 = x + 0.3 * np.array(r2).reshape(-1, 3)
    a = oracle.relax_fire(s, fp, 20000, 1e-3, apos=x1)
    b = oracle.relax_fire(s, fp, 20000, 1e-3, apos=x2)
    assert a["converged"] and b["converged"]
    assert a["fmax"] < 1e-3 and b["fmax"] < 1e-3
    assert abs(a["E"] - b["E"]) < 1e-6


def test_run_md_damped_reaches_convergence(oracle, params):
    s, x = _xylitol_dummy(oracle, params)
    seed, r = _xorshift_stream(0x12345678, 3 * len(x))
    x1 = x + 0.3 * np.array(r).reshape(-1, 3)
    steps, xf, vf = oracle.run_md(s, 20000, 0.01, 1e-3, 1e18, 0.1, apos=x1)
    assert steps < 20000
    f, E = oracle.eval_forces(s, xf)
    assert (f * f).sum() < 1e-6 * 1.5      # sum |F|^2 < fconv^2 at the step it stopped (one move earlier)


def test_pentacene_fire_returns_to_plane(oracle, pentacene):
    """relax_pentacene_uff.rs:118-200,365-397: with inversions off the in-plane
    FIRE relax ends with fmax<0.5 and z_rms<0.05; after an out-of-plane
    distortion FIRE with inversions on returns to z_rms < 0.05."""
    s = oracle.OracleSystem(**{**pentacene.__dict__, "_keep": []})
    inv_saved = s.inv_params.copy()
    s.inv_params = np.tile([0.0, 1.0, -1.0, 0.0], (len(s.inversions), 1))
    fp = oracle.FireParams(dt0=0.001, dt_max=0.05)
    r = oracle.relax_fire(s, fp, 5000, 1e-3)
    zr = np.sqrt((r["apos"][:, 2] ** 2).sum()) / math.sqrt(len(s.elems))
    assert r["fmax"] < 0.5 and zr < 0.05
    s.inv_params = inv_saved
    x = r["apos"].copy()
    i = np.arange(len(x))
    x[:, 2] += 0.1 * (np.sin(i * 0.7) + 0.3 * np.cos(i * 1.3)) / 1.3
    r2 = oracle.relax_fire(s, fp, 5000, 1e-3, apos=x)
    zr2 = np.sqrt((r2["apos"][:, 2] ** 2).sum()) / math.sqrt(len(s.elems))
    assert zr2 < 0.05


def test_system_from_terms_any_order(oracle, params):
    """A system built from given term lists (Uff::from_topology) in shuffled
    order has the parameters of the geometry-built one per term, and the
    same forces and energies to FP64 rounding."""
    from surfmol_amd import workloads as W
    m = W.ptcda_like()
    a = oracle.system_from_geometry(m.elems, m.apos, params, charges=m.charges, bonds=m.bonds)
    rng = np.random.default_rng(3)
    ang = a.angles[rng.permutation(len(a.angles))]
    dih = a.dihedrals[rng.permutation(len(a.dihedrals))]
    inv = a.inversions[rng.permutation(len(a.inversions))]
    b = oracle.system_from_terms(a.types, m.apos, a.bonds, ang, dih, inv, params, charges=m.charges)
    pa = {tuple(d): tuple(p) for d, p in zip(a.dihedrals, a.dih_params)}
    assert all(pa[tuple(d)] == tuple(p) for d, p in zip(b.dihedrals, b.dih_params))
    fa, ea = oracle.eval_forces(a)
    fb, eb = oracle.eval_forces(b)
    assert np.abs(fa - fb).max() < 1e-12 * np.abs(fa).max()
    assert np.abs(ea - eb).max() < 1e-12 * np.abs(ea).max()


def test_gridff_build_nodes_equals_full_build(oracle, params):
    from surfmol_amd import workloads as W
    from surface_setup import slab_reqs
    el, pos, q = W.rocksalt_slab(6.0, 6.0)
    reqs = slab_reqs(params, oracle, el, q)
    ns, lvec = (8, 6, 5), [6.0, 0, 0, 0, 6.0, 0, 0, 0, 2.5]
    E = oracle.gridff_build(pos, reqs, lvec, ns, (0.0, 0.0, 1.0), (1, 1, 0), 1.5, 1.0)
    nodes = np.array([[0, 0, 0], [7, 5, 4], [3, 2, 1], [5, 0, 3]])
    En = oracle.gridff_build_nodes(pos, reqs, lvec, ns, (0.0, 0.0, 1.0), (1, 1, 0), 1.5, 1.0, nodes)
    assert np.array_equal(En, E[nodes[:, 0], nodes[:, 1], nodes[:, 2]])


def test_force_scale(oracle, params):
    """The force scale of the H2 known answer is |F| = 20 on each atom; in
    general it bounds the net force (triangle inequality) and is additive
    over the terms."""
    s = _two_atom_system(oracle)
    assert oracle.force_scale(s) == pytest.approx([20.0, 20.0], abs=1e-12)
    from surfmol_amd import workloads as W
    m = W.ptcda_like()
    o = oracle.system_from_geometry(m.elems, m.apos, params, charges=m.charges, bonds=m.bonds)
    f, _ = oracle.eval_forces(o)
    sc = oracle.force_scale(o)
    assert (sc >= np.linalg.norm(f, axis=1) * (1 - 1e-12)).all()
    o2 = oracle.system_from_geometry(m.elems, m.apos, params, charges=m.charges, bonds=m.bonds, nb_mode=oracle.NB_OFF)
    sc2 = oracle.force_scale(o2)
    assert (sc > sc2).all()
