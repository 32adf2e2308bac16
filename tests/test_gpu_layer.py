"""Coupled layer on the GPU: intermolecular pairs within a cutoff through the
device cell list, one FIRE state for all molecules.

Checked against the FP64 oracle of the same layer (pairs of different
molecules within the cutoff, brute force): per-atom forces within 1e-5 of
max|F| (the cutoff test runs in FP32 on the device; a pair within ~1e-6 A of
the cutoff could fall on the other side), energies within 1e-6 eV/atom; the
relaxed layer is an FP64 minimum; repeated relaxations are bit-identical.
Both the reference's hard cutoff and the smooth switch (energy times
S = 1 - 3x^2 + 2x^3 over the last w below the cutoff) are checked.

Relaxations run on both of the layer's paths: one cooperative launch
(Verlet lists, grid barriers; the default where the layer fits the chip) and
the step launches (force launch, decision, cell list).  Along a trajectory
both follow the FP64 oracle's FIRE steps (positions within 5e-5 A after 1,
7 and 37 steps).
"""
import numpy as np
import pytest

from conftest import rel_force_err
from surface_setup import ALPHA, RDAMP, fire_params_internal, oracle_grid, small_grid_spec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def surf(ctx, oracle, params):
    import surfmol_amd as S
    spec = small_grid_spec()
    E, C, (el, pos, q, reqs) = oracle_grid(oracle, params, spec)
    gs = S.GridSpec(spec["nx"], spec["ny"], spec["nz"], spec["p0"], spec["step"], ALPHA, RDAMP)
    g = S.GridFF.build(ctx, gs, pos, reqs, npbc=2)
    return dict(spec=spec, grid=g, C=g.coefficients().astype(np.float64))


def _layer(ctx, surf, oracle, params, nx, ny, rc, spacing=11.0, seed=7, surface=True, w=0.0):
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    mol, X = W.layer_lattice(nx, ny, spacing=spacing, seed=seed)
    hm = S.Molecule.build(mol.elems, X[0], mol.charges, bonds=mol.bonds)
    b = S.Batch(ctx, [hm], np.zeros(len(X), dtype=np.int32), np.concatenate(X), grid=surf["grid"], rdamp=RDAMP,
                surface=surface, inter_cutoff=rc, inter_switch=w)
    spec = surf["spec"]
    kw = {}
    if surface:
        kw = dict(grid=surf["C"], grid_ns=(spec["nx"], spec["ny"], spec["nz"]), grid_p0=spec["p0"],
                  grid_step=spec["step"], alpha_morse=spec["alpha"])
    osys = oracle.layer_from_molecules([mol] * len(X), X, params, rc, rdamp=RDAMP, rswitch_inter=w, **kw)
    return mol, X, b, osys


@pytest.mark.parametrize("rc,w", [(6.0, 0.0), (12.0, 0.0), (12.0, 2.0)])
def test_layer_eval_forces(ctx, surf, oracle, params, rc, w):
    mol, X, b, osys = _layer(ctx, surf, oracle, params, 4, 3, rc, w=w)
    F, E = b.eval_forces()
    Xd = b.positions()
    fo, eo = oracle.eval_forces(osys, Xd)
    err = rel_force_err(F, fo)
    print("rc", rc, "force err", err)
    assert err < 1e-5
    n = len(Xd)
    # per-molecule energies add up to the layer's (intermolecular pairs
    # counted once at each end, as getNonBond counts every pair twice)
    dE = np.abs(E.sum(axis=0) - eo)
    assert (dE / n < 1e-6).all(), (E.sum(axis=0), eo)


def test_layer_many_cells(ctx, surf, oracle, params):
    """Two groups of molecules 700 A apart with a 6 A cutoff: a grid of
    ~120 x 120 cells, more than the cell build counts in LDS (12 288), so
    its global-count path runs; forces and energies as the FP64 oracle."""
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    mol, X = W.layer_lattice(2, 2, spacing=11.0, seed=7)
    X = list(X) + [x + np.array([700.0, 700.0, 0.0]) for x in X]
    assert ((700.0 + 12.0) / 6.0) ** 2 > 12288
    hm = S.Molecule.build(mol.elems, X[0], mol.charges, bonds=mol.bonds)
    b = S.Batch(ctx, [hm], np.zeros(len(X), dtype=np.int32), np.concatenate(X), grid=surf["grid"], rdamp=RDAMP,
                surface=False, inter_cutoff=6.0)
    osys = oracle.layer_from_molecules([mol] * len(X), X, params, 6.0, rdamp=RDAMP)
    F, E = b.eval_forces()
    Xd = b.positions()
    fo, eo = oracle.eval_forces(osys, Xd)
    assert rel_force_err(F, fo) < 1e-5
    assert (np.abs(E.sum(axis=0) - eo) / len(Xd) < 1e-6).all(), (E.sum(axis=0), eo)


def test_layer_without_neighbours_is_independent(ctx, surf, oracle, params):
    """Molecules 40 A apart: the layer equals the independent batch."""
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    mol, X = W.layer_lattice(2, 2, spacing=40.0, seed=9)
    hm = S.Molecule.build(mol.elems, X[0], mol.charges, bonds=mol.bonds)
    args = (ctx, [hm], np.zeros(4, dtype=np.int32), np.concatenate(X))
    bl = S.Batch(*args, grid=surf["grid"], rdamp=RDAMP, inter_cutoff=12.0)
    bi = S.Batch(*args, grid=surf["grid"], rdamp=RDAMP)
    Fl, El = bl.eval_forces()
    Fi, Ei = bi.eval_forces()
    assert np.array_equal(Fl, Fi)
    assert np.abs(El - Ei).max() < 1e-12


POS_ATOL = 5e-5


@pytest.mark.parametrize("persistent", [True, False])
def test_layer_relax_is_fp64_minimum(ctx, surf, oracle, params, persistent):
    import surfmol_amd as S
    # the smooth switch: with a hard cutoff, pairs crossing it make the
    # forces jump by ~4e-3 eV/A and FIRE may stall above Fmax = 1e-3
    mol, X, b, osys = _layer(ctx, surf, oracle, params, 3, 3, 12.0, w=2.0)
    b.set_layer_persistent(persistent)
    fp = fire_params_internal()
    res = b.relax_batched(S.FireParams(**fp), 3000, 1e-3)
    path = b.layer_path()
    print("path", path)
    assert path["persistent"] == persistent and path["fallback"] is None
    assert all(r["status"] == S.SMR_SYS_CONVERGED for r in res), res[0]
    assert len({r["steps"] for r in res}) == 1          # one FIRE state
    Xf = b.positions()
    fo, eo = oracle.eval_forces(osys, Xf)
    n = len(Xf)
    assert np.linalg.norm(fo, axis=1).max() < 2e-3
    Etot = sum(r["energy"] for r in res)
    assert abs(eo.sum() - Etot) / n < 1e-6, (eo.sum(), Etot)
    rf = oracle.relax_fire(osys, oracle.FireParams(**fp), 3000, 1e-3, apos=Xf)
    assert rf["converged"] and abs(rf["E"] - Etot) / n < 1e-6, (rf["E"], Etot)
    print("layer steps", res[0]["steps"], "E", Etot)


def test_layer_relax_is_reproducible(ctx, surf, oracle, params):
    """Repeated relaxations, replayed from the captured graph or issued
    directly, are bit-identical; so are repeated one-launch relaxations."""
    import surfmol_amd as S
    mol, X, b, osys = _layer(ctx, surf, oracle, params, 4, 4, 12.0, w=2.0)
    fp = S.FireParams(**fire_params_internal())
    p1 = b.relax_batched(fp, 600, 1e-3)
    xp = b.positions()
    assert b.layer_path()["persistent"]
    b.reset()
    p2 = b.relax_batched(fp, 600, 1e-3)
    assert np.array_equal(b.positions(), xp)
    assert [r["energy"] for r in p1] == [r["energy"] for r in p2] and p1[0]["steps"] == p2[0]["steps"]
    b.set_layer_persistent(False)
    b.reset()
    r1 = b.relax_batched(fp, 600, 1e-3)
    x1 = b.positions()
    for graph in (True, False, True):
        b.set_graph(graph)
        b.reset()
        b.eval_forces()
        r2 = b.relax_batched(fp, 600, 1e-3)
        assert np.array_equal(b.positions(), x1), graph
        assert [r["energy"] for r in r1] == [r["energy"] for r in r2]
        assert r1[0]["steps"] == r2[0]["steps"]


# ---------------------------------------------------------------------------
# lateral periodic box (getNonBond's PBC branch, opencl/UFF.cl:1154-1182,
# culled at the cutoff): device against the FP64 oracle's periodic layer
# ---------------------------------------------------------------------------
def _pbc_layer(ctx, surf, oracle, params, nx, ny, spacing, rc, shift=(0.0, 0.0), w=0.0, seed=11):
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    mol, X = W.layer_lattice(nx, ny, spacing=spacing, seed=seed)
    X = [x + np.array([shift[0], shift[1], 0.0]) for x in X]
    box = (nx * spacing, ny * spacing)
    hm = S.Molecule.build(mol.elems, X[0], mol.charges, bonds=mol.bonds)
    b = S.Batch(ctx, [hm], np.zeros(len(X), dtype=np.int32), np.concatenate(X), grid=surf["grid"], rdamp=RDAMP,
                inter_cutoff=rc, inter_switch=w, pbc=box)
    spec = surf["spec"]
    osys = oracle.layer_from_molecules([mol] * len(X), X, params, rc, rdamp=RDAMP, rswitch_inter=w, pbc=box,
                                       grid=surf["C"], grid_ns=(spec["nx"], spec["ny"], spec["nz"]),
                                       grid_p0=spec["p0"], grid_step=spec["step"], alpha_morse=spec["alpha"])
    return mol, X, box, b, osys


@pytest.mark.parametrize("rc,w,shift", [(9.0, 0.0, (0.0, 0.0)), (12.0, 2.0, (-5.3, 7.1)), (12.0, 0.0, (-9.0, -2.0))])
def test_pbc_layer_eval_forces(ctx, surf, oracle, params, rc, w, shift):
    """4 x 4 molecules at 12 A in a 48 x 48 A box (four GridFF periods); the
    shifted lattices put molecules across the box edges."""
    mol, X, box, b, osys = _pbc_layer(ctx, surf, oracle, params, 4, 4, 12.0, rc, shift=shift, w=w)
    F, E = b.eval_forces()
    Xd = b.positions()
    fo, eo = oracle.eval_forces(osys, Xd)
    err = rel_force_err(F, fo)
    print("pbc rc", rc, "shift", shift, "force err", err)
    assert err < 1e-5
    n = len(Xd)
    assert (np.abs(E.sum(axis=0) - eo) / n < 1e-6).all(), (E.sum(axis=0), eo)
    # the images matter: the open layer's forces differ
    import surfmol_amd as S
    bo = S.Batch(ctx, [b.molecules[0]], np.zeros(len(X), dtype=np.int32), np.concatenate(X), grid=surf["grid"],
                 rdamp=RDAMP, inter_cutoff=rc, inter_switch=w)
    Fo, _ = bo.eval_forces()
    assert np.abs(Fo - F).max() > 1e-3


@pytest.mark.parametrize("persistent", [True, False])
def test_pbc_layer_relax_is_fp64_minimum(ctx, surf, oracle, params, persistent):
    import surfmol_amd as S
    mol, X, box, b, osys = _pbc_layer(ctx, surf, oracle, params, 3, 3, 16.0, 12.0, shift=(-4.0, 3.0), w=2.0)
    b.set_layer_persistent(persistent)
    fp = fire_params_internal()
    res = b.relax_batched(S.FireParams(**fp), 4000, 1e-3)
    assert b.layer_path()["persistent"] == persistent
    assert all(r["status"] == S.SMR_SYS_CONVERGED for r in res), res[0]
    Xf = b.positions()
    fo, eo = oracle.eval_forces(osys, Xf)
    n = len(Xf)
    Etot = sum(r["energy"] for r in res)
    assert abs(eo.sum() - Etot) / n < 1e-6, (eo.sum(), Etot)
    rf = oracle.relax_fire(osys, oracle.FireParams(**fp), 4000, 1e-3, apos=Xf)
    assert rf["converged"] and abs(rf["E"] - Etot) / n < 1e-6, (rf["E"], Etot)
    print("pbc layer steps", res[0]["steps"], "E", Etot)


def test_pbc_box_checks(ctx, surf):
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    mol, X = W.layer_lattice(2, 2, spacing=12.0, seed=3)
    hm = S.Molecule.build(mol.elems, X[0], mol.charges, bonds=mol.bonds)
    args = (ctx, [hm], np.zeros(4, dtype=np.int32), np.concatenate(X))
    for box, rc in (((24.0, 24.0), 12.0),      # fewer than 3 cells of the cutoff
                    ((50.0, 48.0), 12.0),      # not a whole number of GridFF periods (12 A)
                    ((48.0, 48.0), 0.0)):      # a box without a coupled layer
        with pytest.raises(S.SmrError):
            S.Batch(*args, grid=surf["grid"], rdamp=RDAMP, inter_cutoff=rc, pbc=box)


# ---------------------------------------------------------------------------
# the two paths of a layer relaxation along a trajectory
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("persistent", [True, False])
@pytest.mark.parametrize("k", [1, 7, 37])
def test_layer_trajectory_matches_oracle(ctx, surf, oracle, params, persistent, k):
    """k FIRE steps of the layer (one shared state; the n_min transition at
    step 6) from the same FP32 start as the oracle's FP64 FIRE."""
    import surfmol_amd as S
    mol, X, b, osys = _layer(ctx, surf, oracle, params, 4, 3, 12.0, w=2.0)
    b.set_layer_persistent(persistent)
    fp = fire_params_internal()
    x0 = b.positions()
    res = b.relax_batched(S.FireParams(**fp), k, 1e-9)
    assert b.layer_path()["persistent"] == persistent
    assert res[0]["steps"] == k
    r = oracle.relax_fire(osys, oracle.FireParams(**fp), k, 1e-9, apos=x0)
    d = np.abs(b.positions() - r["apos"]).max()
    print("persistent", persistent, "k", k, "max position deviation", d)
    assert d < POS_ATOL


@pytest.mark.parametrize("skin", [0.05, 2.0])
def test_one_launch_list_rebuilds(ctx, surf, oracle, params, skin):
    """A thin skin rebuilds the Verlet lists on many steps, a thick one on
    few: both reach the FP64 minimum of the periodic layer."""
    import surfmol_amd as S
    mol, X, box, b, osys = _pbc_layer(ctx, surf, oracle, params, 3, 3, 16.0, 12.0, shift=(-4.0, 3.0), w=2.0)
    b.set_layer_persistent(True, skin)
    fp = fire_params_internal()
    res = b.relax_batched(S.FireParams(**fp), 4000, 1e-3)
    path = b.layer_path()
    print("skin", skin, "steps", res[0]["steps"], path)
    assert path["persistent"] and all(r["status"] == S.SMR_SYS_CONVERGED for r in res)
    if skin < 0.1:
        assert path["rebuilds"] > 20
    Xf = b.positions()
    fo, eo = oracle.eval_forces(osys, Xf)
    Etot = sum(r["energy"] for r in res)
    n = len(Xf)
    assert abs(eo.sum() - Etot) / n < 1e-6
    rf = oracle.relax_fire(osys, oracle.FireParams(**fp), 4000, 1e-3, apos=Xf)
    assert rf["converged"] and abs(rf["E"] - Etot) / n < 1e-6, (rf["E"], Etot)


def test_one_launch_overflow_falls_back(ctx, surf, oracle, params):
    """A 9 A skin stages more neighbour atoms than a molecule's staging area
    holds: the launch stops without writing anything back and the step
    launches relax the layer from the same start (the same minimum)."""
    import surfmol_amd as S
    mol, X, box, b, osys = _pbc_layer(ctx, surf, oracle, params, 4, 4, 12.0, 12.0, w=2.0)
    b.set_layer_persistent(True, 9.0)
    fp = fire_params_internal()
    res = b.relax_batched(S.FireParams(**fp), 200, 1e-3)
    path = b.layer_path()
    print(path)
    assert not path["persistent"] and path["fallback"] is not None and "overflow" in path["fallback"]
    x1 = b.positions()
    b.set_layer_persistent(False)
    b.reset()
    res2 = b.relax_batched(S.FireParams(**fp), 200, 1e-3)
    assert np.array_equal(b.positions(), x1)
    assert [r["energy"] for r in res] == [r["energy"] for r in res2]
