"""Synthetic workload generators: composition, geometry and determinism."""
import math

import numpy as np

from surfmol_amd import workloads as W


def test_ptcda_like_geometry():
    m = W.ptcda_like(jitter=False)
    assert m.natoms == 38
    assert (m.elems.count("C"), m.elems.count("H"), m.elems.count("O")) == (24, 8, 6)
    L = np.linalg.norm(m.apos[m.bonds[:, 1]] - m.apos[m.bonds[:, 0]], axis=1)
    kinds = sorted({(m.elems[a], m.elems[b]) for a, b in m.bonds})
    for (a, b), l in zip(m.bonds, L):
        pair = {m.elems[a], m.elems[b]}
        if "H" in pair:
            assert abs(l - W.B_CH) < 1e-9
        elif pair == {"C", "O"} and (sum(1 for x in m.bonds.flatten() if x == b) == 1 or
                                      sum(1 for x in m.bonds.flatten() if x == a) == 1):
            assert abs(l - W.B_CO) < 1e-9
        else:
            assert abs(l - W.B_RING) < 1e-9
    assert np.abs(m.apos[:, 2]).max() == 0.0


def test_ptcda_like_jitter_and_charges_are_seeded():
    a, b = W.ptcda_like(7), W.ptcda_like(7)
    assert np.array_equal(a.apos, b.apos) and np.array_equal(a.charges, b.charges)
    assert abs(a.charges.sum()) < 1e-12
    assert 0.02 < np.std(a.apos[:, 2]) < 0.1


def test_uneven_sizes_in_range():
    mols = W.uneven_batch(64, seed=11)
    n = np.array([m.natoms for m in mols])
    assert n.min() >= 17 and n.max() <= 420
    assert np.median(n) < 140


def test_slab_is_neutral_and_commensurate():
    el, pos, q = W.rocksalt_slab(24.0, 24.0)
    assert len(el) == 2 * 8 * 8
    assert abs(q.sum()) < 1e-12
    assert set(np.round(pos[:, 2], 9)) == {0.0, -3.0}


def test_fs_units():
    assert math.isclose(W.fs_to_internal(W.FS_PER_TIME_UNIT), 1.0)


def test_polyhex_has_no_clashes():
    """Fused-ring molecules: every pair of lattice neighbours is bonded (the
    induced honeycomb subgraph), coves carry no substituents, so no two
    non-bonded atoms come closer than 1.3 A; every atom has 1..3 bonds."""
    from scipy.spatial.distance import pdist, squareform
    mols = W.uneven_batch(96, seed=W.SEED_BASE + 9)
    for m in mols:
        D = squareform(pdist(m.apos))
        np.fill_diagonal(D, np.inf)
        bonded = np.zeros_like(D, dtype=bool)
        bonded[m.bonds[:, 0], m.bonds[:, 1]] = True
        bonded[m.bonds[:, 1], m.bonds[:, 0]] = True
        assert D[~bonded].min() > 1.3
        deg = np.bincount(m.bonds.flatten(), minlength=m.natoms)
        assert deg.min() >= 1 and deg.max() <= 3
