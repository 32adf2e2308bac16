"""The uniform batch's full-size surface on the GPU: the 24 x 24 x 20 A
rock-salt cell at 0.1 A (240 x 240 x 200 nodes, 11.5 M nodes, 184 MB of
FP32 coefficients), with 16 replicas of the 38-atom PTCDA-like molecule in
the batch's seeded poses.

* Reference grid: device FP64 build at 400 random nodes against the oracle's
  per-node sum (1e-10 relative).
* Fit: S C = E over the whole grid (the oracle applies the B-spline node
  operator to the downloaded FP32 coefficients), 1e-6 of max |E|.
* Forces and energies (1e-6 eV/atom) against the oracle sampling the same
  coefficients, at the start (1e-5 of max_i|F_i|) and after 200 steps (1e-5
  of each atom's sum of contribution magnitudes).
* The relaxed states are FP64 minima.
"""
import numpy as np
import pytest

from conftest import rel_force_err, scaled_force_err
from surface_setup import ALPHA, RDAMP, fire_params_internal, oracle_system, ptcda_replicas, slab_reqs, \
    small_grid_spec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full(ctx, oracle, params):
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    spec = small_grid_spec(L=24.0, step=0.1, zlen=20.0)
    assert (spec["nx"], spec["ny"], spec["nz"]) == (240, 240, 200)
    el, pos, q = W.rocksalt_slab(spec["L"], spec["L"])
    reqs = slab_reqs(params, oracle, el, q)
    gs = S.GridSpec(spec["nx"], spec["ny"], spec["nz"], spec["p0"], spec["step"], ALPHA, RDAMP)
    g = S.GridFF.build(ctx, gs, pos, reqs, npbc=2, keep_reference=True)
    return dict(spec=spec, grid=g, C=g.coefficients().astype(np.float64), sub_pos=pos, sub_reqs=reqs)


def test_full_grid_build_at_sampled_nodes(full, oracle):
    spec = full["spec"]
    ns = (spec["nx"], spec["ny"], spec["nz"])
    rng = np.random.Generator(np.random.PCG64(11))
    nodes = np.stack([rng.integers(0, n, 400) for n in ns], axis=1)
    lvec = [spec["L"], 0, 0, 0, spec["L"], 0, 0, 0, spec["Lz"]]
    Eo = oracle.gridff_build_nodes(full["sub_pos"], full["sub_reqs"], lvec, ns, spec["p0"], (2, 2, 0), ALPHA, RDAMP,
                                   nodes)
    Eg = full["grid"].reference_values()[nodes[:, 0], nodes[:, 1], nodes[:, 2]]
    err = np.abs(Eg - Eo) / np.maximum(1.0, np.abs(Eo))
    assert err.max() < 1e-10, err.max()


def test_full_grid_fit_reproduces_nodes(full, oracle):
    spec = full["spec"]
    ns = (spec["nx"], spec["ny"], spec["nz"])
    E = full["grid"].reference_values()
    SC = oracle.bspline_apply(full["C"], ns)
    rel = np.abs(SC[..., :3] - E[..., :3]).max() / np.abs(E[..., :3]).max()
    print("fit residual at the nodes, relative to max |E|:", rel)
    assert rel < 1e-6
    assert (full["C"][..., 3] == 0).all()


def _batch(ctx, full, mol, X):
    import surfmol_amd as S
    hm = S.Molecule.build(mol.elems, X[0], mol.charges, bonds=mol.bonds)
    return S.Batch(ctx, [hm], np.zeros(len(X), dtype=np.int32), np.concatenate(X), grid=full["grid"], rdamp=RDAMP)


def _check_forces(b, oracle, params, full, mol, nsys, start):
    """At the start: max_i|dF_i| / max_i|F_i|; along the trajectory: per atom
    against the sum of the magnitudes of its contributions."""
    F, E = b.eval_forces()
    Xd = b.positions()
    worst = 0.0
    for s in range(nsys):
        sl = slice(b.offsets[s], b.offsets[s + 1])
        osys = oracle_system(oracle, params, mol, Xd[sl], full["C"], full["spec"])
        fo, eo = oracle.eval_forces(osys)
        err = rel_force_err(F[sl], fo) if start else scaled_force_err(F[sl], fo, oracle.force_scale(osys))
        worst = max(worst, err)
        assert err < 1e-5, (s, err)
        assert (np.abs(E[s] - eo) / mol.natoms < 1e-6).all(), (s, E[s], eo)
    return worst


def test_full_grid_forces_and_relaxation(ctx, full, oracle, params):
    import surfmol_amd as S
    mol, X = ptcda_replicas(16, full["spec"]["L"])
    b = _batch(ctx, full, mol, X)
    w0 = _check_forces(b, oracle, params, full, mol, 16, True)
    fp = S.FireParams(**fire_params_internal())
    b.relax_batched(fp, 200, 1e-3)
    w1 = _check_forces(b, oracle, params, full, mol, 16, False)
    print(f"worst force error: start {w0:.2e}, after 200 steps {w1:.2e}")
    b.reset()
    res = b.relax_batched(fp, 3000, 1e-3)
    Xf = b.positions()
    for s, r in enumerate(res):
        assert r["status"] == S.SMR_SYS_CONVERGED, (s, r)
        sl = slice(b.offsets[s], b.offsets[s + 1])
        osys = oracle_system(oracle, params, mol, X[s], full["C"], full["spec"])
        fo, eo = oracle.eval_forces(osys, Xf[sl])
        assert abs(eo.sum() - r["energy"]) / mol.natoms < 1e-6
        rf = oracle.relax_fire(osys, oracle.FireParams(**fire_params_internal()), 3000, 1e-3, apos=Xf[sl])
        assert rf["converged"] and abs(rf["E"] - r["energy"]) / mol.natoms < 1e-6, (s, rf["E"], r["energy"])


def test_full_uniform_batch(ctx, full, oracle, params):
    """The uniform batch at its full size: 4 096 replicas on the 240 x 240 x
    200 grid, FIRE to Fmax < 1e-3 eV/A within 2 000 steps, the default
    schedule (energy order, re-ranked after 400 steps).  Properties that hold
    at any size: every system converged or stopped at the cap, none
    non-finite; a repeated relaxation, the batch order in one launch and a
    sample of systems relaxed in a batch of their own give the same bits; the
    sample's reported energies are the FP64 energies of its device positions
    and FP64 FIRE from there moves the energy by < 1e-6 eV/atom."""
    import surfmol_amd as S
    nsys, max_steps = 4096, 2000
    mol, X = ptcda_replicas(nsys, full["spec"]["L"])
    b = _batch(ctx, full, mol, X)
    fp = S.FireParams(**fire_params_internal())
    res = b.relax_batched(fp, max_steps, 1e-3)
    Xf = b.positions().copy()
    st = res["status"]
    assert set(np.unique(st)) <= {S.SMR_SYS_CONVERGED, S.SMR_SYS_MAXSTEPS}, np.unique(st)
    conv = st == S.SMR_SYS_CONVERGED
    assert conv.sum() >= nsys - 16, conv.sum()
    assert (res["fmax"][conv] < 1e-3).all()
    assert (res["steps"][~conv] == max_steps).all()
    assert np.isfinite(res["energy"]).all()
    print(f"converged {conv.sum()} of {nsys}, steps mean {res['steps'].mean():.1f} max {res['steps'].max()}")
    # repeated, and in the batch order in one launch
    b.reset()
    res2 = b.relax_batched(fp, max_steps, 1e-3)
    assert np.array_equal(b.positions(), Xf) and np.array_equal(res2["steps"], res["steps"])
    b.set_order(False)
    b.set_rerank(0)
    b.reset()
    res3 = b.relax_batched(fp, max_steps, 1e-3)
    assert np.array_equal(b.positions(), Xf) and np.array_equal(res3["energy"], res["energy"])
    # a sample alone: the longest relaxations and a spread of others
    sample = sorted(set(np.argsort(-res["steps"])[:3].tolist() + list(range(5, nsys, 821))))
    bs = _batch(ctx, full, mol, [X[s] for s in sample])
    rs = bs.relax_batched(fp, max_steps, 1e-3)
    Xs = bs.positions()
    for k, s in enumerate(sample):
        sl, slb = slice(b.offsets[s], b.offsets[s + 1]), slice(bs.offsets[k], bs.offsets[k + 1])
        assert np.array_equal(Xs[slb], Xf[sl]) and rs["steps"][k] == res["steps"][s], s
        osys = oracle_system(oracle, params, mol, X[s], full["C"], full["spec"])
        fo, eo = oracle.eval_forces(osys, Xf[sl])
        assert abs(eo.sum() - res["energy"][s]) / mol.natoms < 1e-6, (s, eo.sum(), res["energy"][s])
        if conv[s]:
            rf = oracle.relax_fire(osys, oracle.FireParams(**fire_params_internal()), 3000, 1e-3, apos=Xf[sl])
            assert rf["converged"] and abs(rf["E"] - res["energy"][s]) / mol.natoms < 1e-6, (s, rf["E"])
