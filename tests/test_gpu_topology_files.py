"""Topologies read from the reference's file formats, relaxed on the GPU.

Uff::from_topology (molff/src/uff.rs:238-240) keeps the file's term lists in
the file's order -- the reference's dihedral order comes from a HashSet
(moltopo/src/topology.rs:141-158) -- and the a2f gather (uff.rs:317-382)
sums each atom's force pieces in that order.  A fused-ring molecule with
N_3, N_R and O_3 atoms is written with its angles, dihedrals and inversions
in a random order, as an UFFTOPO binary (by an independent struct-based
writer of buildff's layout) and as TopologyData JSON; the library reads
each file and the device forces, energies and relaxed state are checked
against the oracle built from the same lists in the same order
(oracle.system_from_terms).
"""
import json

import numpy as np
import pytest

from conftest import rel_force_err
from surface_setup import ALPHA, RDAMP, fire_params_internal, oracle_grid, small_grid_spec
from test_topology_io import _read_ufftopo, _write_ufftopo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def surf(ctx, oracle, params):
    import surfmol_amd as S
    spec = small_grid_spec()
    E, C, (el, pos, q, reqs) = oracle_grid(oracle, params, spec)
    gs = S.GridSpec(spec["nx"], spec["ny"], spec["nz"], spec["p0"], spec["step"], ALPHA, RDAMP)
    g = S.GridFF.build(ctx, gs, pos, reqs, npbc=2)
    return dict(spec=spec, grid=g, C=g.coefficients().astype(np.float64))


def _molecule():
    """A seeded fused-ring molecule that has N_3, N_R and O_3 atoms."""
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(W.SEED_BASE + 77))
    for _ in range(200):
        m = W.random_polyhex(rng, 60)
        types = S.Molecule.build(m.elems, m.apos, m.charges, bonds=m.bonds).types
        if {"N_3", "N_R", "O_3"} <= set(types):
            ang, xy, h = W.uniform_batch_poses(1, L=12.0, hmin=3.5, hmax=3.5, seed=W.SEED_BASE + 78)
            return m, W.place_on_surface(m, ang[0], xy[0], h[0])
    raise AssertionError("no molecule with N_3, N_R and O_3 found")


def _shuffled_terms(hm, seed):
    t = hm.tables()
    rng = np.random.default_rng(seed)
    return (t["types"], t["bonds"], t["ang_atoms"][rng.permutation(len(t["ang_atoms"]))],
            t["dih_atoms"][rng.permutation(len(t["dih_atoms"]))], t["inv_atoms"][rng.permutation(len(t["inv_atoms"]))])


def _oracle_sys(oracle, params, surf, m, x, types, bonds, ang, dih, inv):
    spec = surf["spec"]
    return oracle.system_from_terms(types, x, bonds, ang, dih, inv, params, charges=m.charges, nb_mode=oracle.NB_GPU,
                                    rdamp=RDAMP, grid=surf["C"], grid_ns=(spec["nx"], spec["ny"], spec["nz"]),
                                    grid_p0=spec["p0"], grid_step=spec["step"], alpha_morse=spec["alpha"])


@pytest.mark.parametrize("fmt", ["ufftopo", "json"])
def test_file_topology_relaxes_as_oracle(ctx, surf, oracle, params, tmp_path, fmt):
    import surfmol_amd as S
    m, x = _molecule()
    built = S.Molecule.build(m.elems, x, m.charges, bonds=m.bonds)
    types, bonds, ang, dih, inv = _shuffled_terms(built, 5)
    path = tmp_path / f"mol.{fmt}"
    if fmt == "ufftopo":
        _write_ufftopo(path, x, types, bonds, ang, dih, inv)
        f = _read_ufftopo(path)
        assert f["types"] == types
        hm = S.Molecule.read_ufftopo(path, charges=m.charges)
    else:
        d = {"natoms": len(types), "elements": list(m.elems), "positions": x.tolist(), "bonds": bonds.tolist(),
             "angles": ang.tolist(), "dihedrals": dih.tolist(), "inversions": inv.tolist(), "bond_params": [],
             "angle_params": [], "dihedral_params": [], "inversion_params": [], "atom_params": []}
        path.write_text(json.dumps(d))
        hm = S.Molecule.read_json(path, charges=m.charges)
    t = hm.tables()
    np.testing.assert_array_equal(t["dih_atoms"], dih)     # the file's order is kept
    np.testing.assert_array_equal(t["ang_atoms"], ang)
    np.testing.assert_array_equal(t["inv_atoms"], inv)
    b = S.Batch(ctx, [hm], [0], x, grid=surf["grid"], rdamp=RDAMP)
    X0 = b.positions()
    osys = _oracle_sys(oracle, params, surf, m, X0, types, bonds, ang, dih, inv)
    F, E = b.eval_forces()
    fo, eo = oracle.eval_forces(osys)
    assert rel_force_err(F, fo) < 1e-5
    assert (np.abs(E[0] - eo) / len(types) < 1e-6).all(), (E[0], eo)
    fpd = fire_params_internal()
    r = b.relax_batched(S.FireParams(**fpd), 3000, 1e-3)[0]
    assert r["status"] == S.SMR_SYS_CONVERGED, r
    Xf = b.positions()
    fo, eo = oracle.eval_forces(osys, Xf)
    assert abs(eo.sum() - r["energy"]) / len(types) < 1e-6
    rf = oracle.relax_fire(osys, oracle.FireParams(**fpd), 3000, 1e-3, apos=Xf)
    assert rf["converged"] and abs(rf["E"] - r["energy"]) / len(types) < 1e-6, (rf["E"], r["energy"])
