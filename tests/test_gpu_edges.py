"""Edge sizes and statuses on the GPU: the smallest systems (one atom:
surface only; two atoms: one bond, no pair), the largest supported system
(exactly 1024 atoms), requests beyond what is built, which are errors, and
the step-cap and non-finite statuses (a non-finite system stops alone).

The reference's relax_multi takes systems of any size (relax_multi.rs reads
natoms per system from the topology); this library supports 1..1024 atoms
per system and says so with SMR_ERR_UNSUPPORTED above that.  Forces and
energies are checked against the FP64 oracle with the tolerances of the
other batch tests (forces 1e-5 relative per system, energies 1e-6 eV/atom).
"""
import types

import numpy as np
import pytest

from conftest import rel_force_err
from surface_setup import ALPHA, RDAMP, fire_params_internal, oracle_grid, oracle_system, small_grid_spec

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
E_ATOL_PER_ATOM = 1e-6
MAX_ATOMS = 1024


@pytest.fixture(scope="module")
def surf(ctx, oracle, params):
    import surfmol_amd as S
    spec = small_grid_spec()
    E, C, (el, pos, q, reqs) = oracle_grid(oracle, params, spec)
    gs = S.GridSpec(spec["nx"], spec["ny"], spec["nz"], spec["p0"], spec["step"], ALPHA, RDAMP)
    g = S.GridFF.build(ctx, gs, pos, reqs, npbc=2)
    return dict(spec=spec, grid=g, C=g.coefficients().astype(np.float64))


def _mol(elems, charges, bonds):
    return types.SimpleNamespace(elems=list(elems), charges=np.asarray(charges, dtype=float),
                                 bonds=np.asarray(bonds, dtype=np.int32).reshape(-1, 2), natoms=len(elems))


def _tiny():
    """A lone C, a lone O (charged), CO, and a 22-atom polyhex beside them
    in the same size class."""
    from surfmol_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(0x5EED3001))
    poly = W.random_polyhex(rng, 22)
    ang, xy, h = W.uniform_batch_poses(1, L=12.0, hmin=3.2, hmax=3.2, seed=0x5EED3002)
    mols = [_mol(["C"], [0.05], []), _mol(["O"], [-0.2], []), _mol(["C", "O"], [0.1, -0.1], [[0, 1]]), poly]
    X = [np.array([[1.3, 2.1, 3.4]]), np.array([[7.7, 4.2, 3.1]]), np.array([[3.0, 8.0, 3.3], [4.13, 8.05, 3.25]]),
         W.place_on_surface(poly, ang[0], xy[0], h[0])]
    return mols, X


def _max_size():
    """Exactly 1024 atoms: a fused-ring sheet of about 940 atoms with C atoms
    on a 3.3 A lattice 3.4 A above it, unbonded, filling up the count."""
    from surfmol_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(0x5EED3003))
    poly = W.random_polyhex(rng, MAX_ATOMS)
    assert poly.natoms < MAX_ATOMS
    ang, xy, h = W.uniform_batch_poses(1, L=12.0, hmin=3.3, hmax=3.3, seed=0x5EED3004)
    Xp = W.place_on_surface(poly, ang[0], xy[0], h[0])
    npad = MAX_ATOMS - poly.natoms
    side = int(np.ceil(np.sqrt(npad)))
    c = Xp.mean(axis=0)
    g = (np.arange(side) - (side - 1) / 2) * 3.3
    pad = np.array([[c[0] + gx, c[1] + gy, Xp[:, 2].max() + 3.4] for gx in g for gy in g][:npad])
    m = _mol(list(poly.elems) + ["C"] * npad, np.concatenate([poly.charges, np.zeros(npad)]), poly.bonds)
    return m, np.vstack([Xp, pad])


def _batch(ctx, surf, mols, X):
    import surfmol_amd as S
    hm = [S.Molecule.build(m.elems, X[s], m.charges, bonds=m.bonds) for s, m in enumerate(mols)]
    return S.Batch(ctx, hm, np.arange(len(mols), dtype=np.int32), np.concatenate(X), grid=surf["grid"], rdamp=RDAMP)


def _check_eval(b, mols, oracle, params, surf):
    F, E = b.eval_forces()
    Xd = b.positions()
    for s, m in enumerate(mols):
        sl = slice(b.offsets[s], b.offsets[s + 1])
        osys = oracle_system(oracle, params, m, Xd[sl], surf["C"], surf["spec"])
        fo, eo = oracle.eval_forces(osys)
        assert np.linalg.norm(fo, axis=1).max() > 1e-3, (s, fo)   # a real force to compare
        err = rel_force_err(F[sl], fo)
        assert err < FORCE_RTOL, (s, m.natoms, err)
        dE = np.abs(E[s] - eo)
        assert (dE / m.natoms < E_ATOL_PER_ATOM).all(), (s, m.natoms, dE, eo)


@pytest.mark.parametrize("block", [0, 64, 128, 256])
@pytest.mark.parametrize("pair_sweep", [1, 0])
def test_tiny_systems_eval(ctx, surf, oracle, params, block, pair_sweep):
    """1- and 2-atom systems in every launch shape of their size class: no
    pair (N=1), one pair that is a bond (N=2) and a sweep of 11 (N=22)."""
    mols, X = _tiny()
    b = _batch(ctx, surf, mols, X)
    b.set_launch(block, -1, pair_sweep)
    print(b.plan())
    _check_eval(b, mols, oracle, params, surf)
    b.close()


def test_tiny_systems_relax(ctx, surf, oracle, params):
    """Lone atoms and CO relax on the surface to the FP64 minimum of the
    same start (one adsorption site each: the energies agree closely)."""
    import surfmol_amd as S
    mols, X = _tiny()
    b = _batch(ctx, surf, mols, X)
    fp = fire_params_internal()
    res = b.relax_batched(S.FireParams(**fp), 3000, 1e-3)
    Xf = b.positions()
    for s, (m, r) in enumerate(zip(mols, res)):
        print(s, m.natoms, r)
        assert r["status"] == S.SMR_SYS_CONVERGED, (s, r)
        sl = slice(b.offsets[s], b.offsets[s + 1])
        osys = oracle_system(oracle, params, m, X[s], surf["C"], surf["spec"])
        fo, eo = oracle.eval_forces(osys, Xf[sl])
        assert abs(eo.sum() - r["energy"]) / m.natoms < E_ATOL_PER_ATOM, (s, eo.sum(), r["energy"])
        rf = oracle.relax_fire(osys, oracle.FireParams(**fp), 3000, 1e-3, apos=Xf[sl])
        assert rf["converged"]
        assert abs(rf["E"] - r["energy"]) / m.natoms < E_ATOL_PER_ATOM, (s, rf["E"], r["energy"])
        if m.natoms <= 2:
            ro = oracle.relax_fire(osys, oracle.FireParams(**fp), 3000, 1e-3)
            assert ro["converged"]
            assert abs(ro["E"] - r["energy"]) < 1e-5, (s, ro["E"], r["energy"])
            # the same adsorption site: both stop at Fmax < 1e-3 eV/A, which on
            # a lateral curvature of ~0.1 eV/A^2 leaves ~1e-2 A of play
            assert np.abs(ro["apos"] - Xf[sl]).max() < 5e-2, (s, ro["apos"], Xf[sl])
    b.close()


def _every_size(nmin=2, nmax=64):
    """One system of each size nmin..nmax: a fused-ring molecule (bonds,
    so 1-2 exclusions) topped up to the exact count with unbonded C atoms
    3.4 A above it."""
    from surfmol_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(0x5EED3005))
    mols, X = [], []
    for n in range(nmin, nmax + 1):
        poly = None
        for target in (int(0.75 * n), int(0.6 * n), int(0.45 * n)):
            if target >= 12:
                cand = W.random_polyhex(rng, target)
                if cand.natoms <= n:
                    poly = cand
                    break
        if poly is None:
            # small sizes: bonded C-O pairs 3.3 A apart
            k = n // 2
            el = ["C", "O"] * k
            pos = np.array([[3.3 * q, 0.0, 0.0] if e == "C" else [3.3 * q + 1.13, 0.0, 0.0]
                            for q in range(k) for e in ("C", "O")]).reshape(-1, 3)
            poly = types.SimpleNamespace(elems=el, charges=np.tile([0.1, -0.1], k), bonds=np.array(
                [[2 * q, 2 * q + 1] for q in range(k)], dtype=np.int32).reshape(-1, 2), natoms=2 * k, apos=pos)
            Xp = pos + np.array([4.0, 4.0, 3.3])
        else:
            ang, xy, h = W.uniform_batch_poses(1, L=12.0, hmin=3.3, hmax=3.3, seed=0x5EED3006 + n)
            Xp = W.place_on_surface(poly, ang[0], xy[0], h[0])
        npad = n - poly.natoms
        c = Xp.mean(axis=0)
        side = max(1, int(np.ceil(np.sqrt(npad))))
        g = (np.arange(side) - (side - 1) / 2) * 3.3
        pad = np.array([[c[0] + gx, c[1] + gy, Xp[:, 2].max() + 3.4] for gx in g for gy in g][:npad]).reshape(-1, 3)
        mols.append(_mol(list(poly.elems) + ["C"] * npad, np.concatenate([poly.charges, np.zeros(npad)]), poly.bonds))
        X.append(np.vstack([Xp, pad]))
    return mols, X


def test_sweep_schedule_every_size(ctx, surf, oracle, params):
    """The one-wave pair sweep's lane schedules (one per size up to 43 atoms,
    the wrapping rotation above): every size from 2 to 64 atoms in the
    64-thread shape, forces against the FP64 oracle."""
    mols, X = _every_size()
    assert [m.natoms for m in mols] == list(range(2, 65))
    b = _batch(ctx, surf, mols, X)
    b.set_launch(64, -1, 1)
    print(b.plan())
    _check_eval(b, mols, oracle, params, surf)
    b.close()


@pytest.mark.parametrize("block", [0, 512])
@pytest.mark.parametrize("pair_sweep", [-1, 0])
def test_max_size_system_eval(ctx, surf, oracle, params, block, pair_sweep):
    m, X = _max_size()
    assert m.natoms == MAX_ATOMS
    b = _batch(ctx, surf, [m], [X])
    b.set_launch(block, -1, pair_sweep)
    plan = b.plan()
    print(plan)
    assert plan[0]["max_atoms"] == MAX_ATOMS and plan[0]["cls"] == 4
    _check_eval(b, [m], oracle, params, surf)
    b.close()


def test_max_size_pair_sweep_demand_is_an_error(ctx, surf, oracle, params):
    """The unordered-pair sweep's per-wave buffers for 1024 atoms exceed the
    LDS of every shape: demanding the sweep is an error before anything is
    launched, and the batch works with the automatic choice afterwards."""
    import surfmol_amd as S
    from surfmol_amd import _lib as L
    m, X = _max_size()
    b = _batch(ctx, surf, [m], [X])
    b.set_launch(0, -1, 1)
    with pytest.raises(S.SmrError) as e:
        b.eval_forces()
    assert e.value.code == L.SMR_ERR_UNSUPPORTED and "pair_sweep" in str(e.value)
    b.set_launch(0, -1, -1)
    assert not b.plan()[0]["pair_sweep"]
    _check_eval(b, [m], oracle, params, surf)
    b.close()


def test_max_size_system_in_mix_relaxes_alike(ctx, surf):
    """A 1024-atom system beside small ones: the first 200 steps give the
    same positions as the system alone."""
    import surfmol_amd as S
    m, X = _max_size()
    tiny, Xt = _tiny()
    fp = S.FireParams(**fire_params_internal())
    alone = _batch(ctx, surf, [m], [X])
    ra = alone.relax_batched(fp, 200, 1e-3)[0]
    mix = _batch(ctx, surf, tiny + [m], Xt + [X])
    rm = mix.relax_batched(fp, 200, 1e-3)[-1]
    assert ra["steps"] == rm["steps"] == 200 and ra["energy"] == rm["energy"], (ra, rm)
    assert np.array_equal(alone.positions(), mix.positions()[mix.offsets[-2]:])
    alone.close()
    mix.close()


def test_size_limits_are_errors(ctx, surf):
    import surfmol_amd as S
    from surfmol_amd import _lib as L
    m, X = _max_size()
    big = _mol(m.elems + ["C"], np.append(m.charges, 0.0), m.bonds)
    Xb = np.vstack([X, X[-1] + [0.0, 0.0, 3.3]])
    hm = S.Molecule.build(big.elems, Xb, big.charges, bonds=big.bonds)   # topology itself has no limit
    assert hm.natoms == MAX_ATOMS + 1
    with pytest.raises(S.SmrError) as e:
        S.Batch(ctx, [hm], [0], Xb, grid=surf["grid"], rdamp=RDAMP)
    assert e.value.code == L.SMR_ERR_UNSUPPORTED and "1025" in str(e.value)
    with pytest.raises(S.SmrError) as e:
        S.Batch(ctx, [hm], np.zeros(0, dtype=np.int32), np.zeros((0, 3)), grid=surf["grid"], rdamp=RDAMP)
    assert e.value.code == L.SMR_ERR_INVALID
    small = S.Molecule.build(["C"], np.array([[0.0, 0.0, 3.0]]))
    with pytest.raises(S.SmrError) as e:   # a topology index outside the list
        S.Batch(ctx, [small], [-1], np.array([[0.0, 0.0, 3.0]]), grid=surf["grid"], rdamp=RDAMP)
    assert e.value.code == L.SMR_ERR_INVALID


def test_nonfinite_input_is_an_error(ctx, surf):
    import surfmol_amd as S
    from surfmol_amd import _lib as L
    mols, X = _tiny()
    hm = S.Molecule.build(mols[2].elems, X[2], mols[2].charges, bonds=mols[2].bonds)
    bad = X[2].copy()
    bad[1, 0] = np.nan
    with pytest.raises(S.SmrError) as e:
        S.Batch(ctx, [hm], [0], bad, grid=surf["grid"], rdamp=RDAMP)
    assert e.value.code == L.SMR_ERR_INVALID


def test_step_cap_status(ctx, surf, oracle, params):
    """A relaxation cut at max_steps reports SMR_SYS_MAXSTEPS after exactly
    that many steps, with the FP64 energy of the positions it returns."""
    import surfmol_amd as S
    mols, X = _tiny()
    b = _batch(ctx, surf, mols, X)
    res = b.relax_batched(S.FireParams(**fire_params_internal()), 10, 1e-3)
    Xf = b.positions()
    for s, (m, r) in enumerate(zip(mols, res)):
        assert r["status"] == S.SMR_SYS_MAXSTEPS and r["steps"] == 10, (s, r)
        sl = slice(b.offsets[s], b.offsets[s + 1])
        fo, eo = oracle.eval_forces(oracle_system(oracle, params, m, X[s], surf["C"], surf["spec"]), Xf[sl])
        assert abs(eo.sum() - r["energy"]) / m.natoms < E_ATOL_PER_ATOM, (s, eo.sum(), r["energy"])
    b.close()


def test_nonfinite_system_stops_alone(ctx, surf):
    """Two unbonded atoms on the same spot give a non-finite pair force: that
    system stops with SMR_SYS_NONFINITE after one step, and the other system
    of the batch relaxes bit for bit as in a batch of its own."""
    import surfmol_amd as S
    mols, X = _tiny()
    Xt = np.array([[2.0, 3.0, 3.4], [2.0, 3.0, 3.4]])
    # no bonds: nbonds == 0 with rcut_bond = 0 (no pair is closer than 0)
    twin = S.Molecule.build(["C", "C"], Xt, [0.0, 0.0], rcut_bond=0.0)
    assert twin.counts == (2, 0, 0, 0, 0)
    poly = S.Molecule.build(mols[3].elems, X[3], mols[3].charges, bonds=mols[3].bonds)
    fp = S.FireParams(**fire_params_internal())
    b = S.Batch(ctx, [twin, poly], [0, 1], np.concatenate([Xt, X[3]]), grid=surf["grid"], rdamp=RDAMP)
    res = b.relax_batched(fp, 2000, 1e-3)
    assert res[0]["status"] == S.SMR_SYS_NONFINITE and res[0]["steps"] == 1, res[0]
    a = _batch(ctx, surf, [mols[3]], [X[3]])
    ra = a.relax_batched(fp, 2000, 1e-3)[0]
    assert res[1]["status"] == ra["status"] == S.SMR_SYS_CONVERGED
    assert res[1]["steps"] == ra["steps"] and res[1]["energy"] == ra["energy"], (res[1], ra)
    assert np.array_equal(b.positions()[b.offsets[1]:], a.positions())
    a.close()
    b.close()
