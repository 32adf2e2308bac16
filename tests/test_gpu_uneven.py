"""Uneven batches on the GPU: systems of different sizes share one batch and
are split into size classes (one persistent launch each, concurrent
streams, captured in a HIP graph).

Checked against the FP64 oracle with the same tolerances as the uniform
batch (forces 1e-5 relative per system, energies 1e-6 eV/atom), plus: the
unordered-pair sweep and the ordered pair loop agree, a graph replayed
several times gives the eager result bit for bit, every system relaxes the
same alone and in the mix, and autotuned shapes relax to the same basins.
"""
import numpy as np
import pytest

from conftest import rel_force_err
from surface_setup import ALPHA, RDAMP, fire_params_internal, oracle_grid, oracle_system, small_grid_spec

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
E_ATOL_PER_ATOM = 1e-6

# one or more molecules per size class (<=64, <=128, <=256, <=512, <=1024 atoms)
EVAL_SIZES = [22, 47, 64, 90, 128, 170, 256, 330, 480, 700]
RELAX_SIZES = [24, 40, 58, 75, 110, 150]


@pytest.fixture(scope="module")
def surf(ctx, oracle, params):
    import surfmol_amd as S
    spec = small_grid_spec()
    E, C, (el, pos, q, reqs) = oracle_grid(oracle, params, spec)
    gs = S.GridSpec(spec["nx"], spec["ny"], spec["nz"], spec["p0"], spec["step"], ALPHA, RDAMP)
    g = S.GridFF.build(ctx, gs, pos, reqs, npbc=2)
    return dict(spec=spec, grid=g, C=g.coefficients().astype(np.float64))


def _molecules(sizes, seed):
    from surfmol_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(seed))
    return [W.random_polyhex(rng, n) for n in sizes]


def _poses(mols, L, seed):
    from surfmol_amd import workloads as W
    ang, xy, h = W.uniform_batch_poses(len(mols), L=L, hmin=3.0, hmax=4.5, seed=seed)
    return [W.place_on_surface(m, ang[s], xy[s], h[s]) for s, m in enumerate(mols)]


def _batch(ctx, surf, mols, X, **kw):
    import surfmol_amd as S
    hm = [S.Molecule.build(m.elems, X[s], m.charges, bonds=m.bonds) for s, m in enumerate(mols)]
    return S.Batch(ctx, hm, np.arange(len(mols), dtype=np.int32), np.concatenate(X), grid=surf["grid"], rdamp=RDAMP,
                   **kw)


@pytest.mark.parametrize("block", [0, 64, 128, 512])
@pytest.mark.parametrize("pair_sweep", [1, 0])
def test_uneven_eval_forces(ctx, surf, oracle, params, pair_sweep, block):
    """Every launch shape (64/128/256/512 threads x 1/2/4 atoms per thread,
    pair sweep or ordered pairs) gives the FP64 forces and energies."""
    mols = _molecules(EVAL_SIZES, 0x5EED1001)
    X = _poses(mols, surf["spec"]["L"], 0x5EED1002)
    b = _batch(ctx, surf, mols, X)
    b.set_launch(block, -1, pair_sweep)
    plan = b.plan()
    print([(p["cls"], p["systems"], p["max_atoms"], p["block"], p["apt"], p["lds_bytes"], p["pair_sweep"],
            p["fint_lds"]) for p in plan])
    # every system in exactly one launch (adjacent classes of one shape may share one)
    assert sum(p["systems"] for p in plan) == len(mols)
    assert len({p["cls"] for p in plan}) == len(plan) and {p["cls"] for p in plan} <= {0, 1, 2, 3, 4}
    for p in plan:
        if pair_sweep == 0:
            assert not p["pair_sweep"]
        # a forced width is used wherever a shape of that width fits; 512
        # threads with 2 atoms each stage tables that exceed the LDS of the
        # largest class, which then falls back to 256 x 4
        if block and p["max_atoms"] <= (block if block == 512 else 4 * block):
            assert p["block"] == block
    F, E = b.eval_forces()
    Xd = b.positions()
    for s, m in enumerate(mols):
        sl = slice(b.offsets[s], b.offsets[s + 1])
        osys = oracle_system(oracle, params, m, Xd[sl], surf["C"], surf["spec"])
        fo, eo = oracle.eval_forces(osys)
        err = rel_force_err(F[sl], fo)
        assert err < FORCE_RTOL, (s, m.natoms, err)
        dE = np.abs(E[s] - eo)
        assert (dE / m.natoms < E_ATOL_PER_ATOM).all(), (s, m.natoms, dE, eo)


def test_uneven_relax_matches_fp64(ctx, surf, oracle, params):
    import surfmol_amd as S
    mols = _molecules(RELAX_SIZES, 0x5EED1003)
    X = _poses(mols, surf["spec"]["L"], 0x5EED1004)
    b = _batch(ctx, surf, mols, X)
    fp = fire_params_internal()
    res = b.relax_batched(S.FireParams(**fp), 3000, 1e-3)
    Xf = b.positions()
    # the device's energies at the final positions (FP32 forces, FP64 sums)
    _, Ef = b.eval_forces()
    for s, (m, r) in enumerate(zip(mols, res)):
        assert r["status"] == S.SMR_SYS_CONVERGED, (s, m.natoms, r)
        sl = slice(b.offsets[s], b.offsets[s + 1])
        osys = oracle_system(oracle, params, m, X[s], surf["C"], surf["spec"])
        fo, eo = oracle.eval_forces(osys, Xf[sl])
        assert np.linalg.norm(fo, axis=1).max() < 2e-3, (s, m.natoms)
        # converged energy (of the returned positions) against FP64 there
        assert abs(eo.sum() - r["energy"]) / m.natoms < E_ATOL_PER_ATOM, (s, eo.sum(), r["energy"])
        assert abs(eo.sum() - Ef[s].sum()) / m.natoms < E_ATOL_PER_ATOM, (s, eo.sum(), Ef[s].sum())
        # the device's minimum is an FP64 minimum: FP64 FIRE started there
        # converges without moving the energy by more than 1e-6 eV/atom
        rf = oracle.relax_fire(osys, oracle.FireParams(**fp), 3000, 1e-3, apos=Xf[sl])
        assert rf["converged"]
        assert abs(rf["E"] - r["energy"]) / m.natoms < E_ATOL_PER_ATOM, (s, rf["E"], r["energy"])
        # an FP64 relaxation from the same start ends in the same region of
        # the landscape (flexible molecules have nearby minima: FP32 and FP64
        # trajectories may pick different ones)
        ro = oracle.relax_fire(osys, oracle.FireParams(**fp), 3000, 1e-3)
        assert ro["converged"]
        print(s, m.natoms, r["steps"], ro["steps"], r["energy"], ro["E"])
        assert abs(r["energy"] - ro["E"]) / m.natoms < 1e-3, (s, m.natoms, r["energy"], ro["E"])


def test_uneven_same_alone_and_in_mix(ctx, surf):
    """A system's relaxation does not depend on the other systems of the
    batch (its size class fixes its launch shape)."""
    import surfmol_amd as S
    mols = _molecules([30, 70, 150, 300, 45, 120], 0x5EED1005)
    X = _poses(mols, surf["spec"]["L"], 0x5EED1006)
    fp = S.FireParams(**fire_params_internal())
    mix = _batch(ctx, surf, mols, X).relax_batched(fp, 1500, 1e-3)
    for s in [0, 1, 2, 3]:
        alone = _batch(ctx, surf, mols[s:s + 1], X[s:s + 1]).relax_batched(fp, 1500, 1e-3)
        assert alone[0]["steps"] == mix[s]["steps"], (s, alone[0], mix[s])
        assert alone[0]["energy"] == mix[s]["energy"], (s, alone[0], mix[s])


def test_merged_launch_same_as_alone(ctx, surf):
    """With 256 threads forced, the <= 128- and <= 256-atom classes share one
    persistent launch; each system relaxes bit for bit as in a batch of its
    own (its launch shape is the same either way)."""
    import surfmol_amd as S
    mols = _molecules([90, 120, 150, 200, 250, 100], 0x5EED100B)
    X = _poses(mols, surf["spec"]["L"], 0x5EED100C)
    fp = S.FireParams(**fire_params_internal())
    b = _batch(ctx, surf, mols, X)
    b.set_launch(256, -1, -1)
    plan = b.plan()
    print(plan)
    assert len(plan) == 1 and plan[0]["systems"] == len(mols) and plan[0]["block"] == 256, plan
    mix = b.relax_batched(fp, 1200, 1e-3)
    xm = b.positions()
    for s in [0, 3, 4]:
        a = _batch(ctx, surf, mols[s:s + 1], X[s:s + 1])
        a.set_launch(256, -1, -1)
        ra = a.relax_batched(fp, 1200, 1e-3)[0]
        assert ra["steps"] == mix[s]["steps"] and ra["energy"] == mix[s]["energy"], (s, ra, mix[s])
        assert np.array_equal(a.positions(), xm[b.offsets[s]:b.offsets[s + 1]])
        a.close()
    b.close()


def test_graph_replays_match_eager(ctx, surf):
    """The captured launch sequence, replayed several times with eager work
    and synchronisations between replays, gives the eager result each time."""
    import surfmol_amd as S
    mols = _molecules([26, 80, 200], 0x5EED1007)
    X = _poses(mols, surf["spec"]["L"], 0x5EED1008)
    fp = S.FireParams(**fire_params_internal())
    b = _batch(ctx, surf, mols, X)
    b.set_graph(False)
    ref = b.relax_batched(fp, 400, 1e-3)
    xref = b.positions()
    b.set_graph(True)
    for rep in range(3):
        b.reset()
        F, E = b.eval_forces()          # eager work between replays
        r = b.relax_batched(fp, 400, 1e-3)
        x = b.positions()
        assert np.array_equal(x, xref), rep
        for a, c in zip(ref, r):
            assert a["steps"] == c["steps"] and a["energy"] == c["energy"], (rep, a, c)


def test_autotune_same_basins(ctx, surf):
    import surfmol_amd as S
    mols = _molecules([33, 66, 140, 260], 0x5EED1009)
    X = _poses(mols, surf["spec"]["L"], 0x5EED100A)
    fp = S.FireParams(**fire_params_internal())
    r0 = _batch(ctx, surf, mols, X).relax_batched(fp, 3000, 1e-3)
    # a private context: tuned shapes are remembered per context
    c2 = S.Context(0)
    b = _batch(c2, surf_for(c2, surf), mols, X)
    plan = b.autotune(fp, trial_steps=20)
    print(plan)
    ncls = len({min(k for k in range(5) if m.natoms <= (64 << k)) for m in mols})
    assert 1 <= len(plan) <= ncls and all(p["block"] > 0 for p in plan)
    assert sum(p["systems"] for p in plan) == len(mols)
    r1 = b.relax_batched(fp, 3000, 1e-3)
    for m, a, c in zip(mols, r0, r1):
        print(m.natoms, a["status"], a["steps"], a["energy"], c["status"], c["steps"], c["energy"])
    nconv = 0
    for m, a, c in zip(mols, r0, r1):
        # a system that stalls on a soft mode may run out of steps with
        # either shape; compare the relaxations that converge
        if a["status"] == S.SMR_SYS_CONVERGED and c["status"] == S.SMR_SYS_CONVERGED:
            nconv += 1
            assert abs(a["energy"] - c["energy"]) / m.natoms < 2e-5, (a, c)
    assert nconv >= 3
    b.close()
    c2.close()


def surf_for(ctx, surf):
    """The same grid uploaded into another context."""
    import surfmol_amd as S
    spec = surf["spec"]
    gs = S.GridSpec(spec["nx"], spec["ny"], spec["nz"], spec["p0"], spec["step"], ALPHA, RDAMP)
    return dict(spec=spec, grid=S.GridFF.upload(ctx, gs, surf["grid"].coefficients()), C=surf["C"])
