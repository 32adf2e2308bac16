"""smr_relax_plain_noalloc (the plain method with its kernels' scratch in a
fixed device array) against smr_relax_plain, through the C ABI: two calls in
a row on one GPU, the second on a smaller system, each bit-identical to the
allocating entry point (same kernels, same order of operations) in result
and final positions."""
import ctypes

import numpy as np
import pytest

from surface_setup import ALPHA, RDAMP, fire_params_internal, ptcda_replicas, small_grid_spec

pytestmark = pytest.mark.gpu


def _noalloc(b, isys, fp, max_steps, fconv):
    import surfmol_amd as S
    from surfmol_amd import _lib
    res = np.zeros(1, dtype=S.RESULT_DTYPE)
    S._lib.check(S.lib().smr_relax_plain_noalloc(b._h, int(isys), ctypes.byref(fp.cstruct()), int(max_steps),
                                                 float(fconv), res.ctypes.data_as(ctypes.POINTER(_lib.SmrRelaxResult))))
    return res[0]


def test_plain_noalloc_matches_plain_twice(ctx):
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    spec = small_grid_spec()
    el, pos, q = W.rocksalt_slab(spec["L"], spec["L"])
    gs = S.GridSpec(spec["nx"], spec["ny"], spec["nz"], spec["p0"], spec["step"], ALPHA, RDAMP)
    g = S.GridFF.build(ctx, gs, pos, S.substrate_reqs(el, q), npbc=2)
    fp = S.FireParams(**fire_params_internal())
    mol, X = ptcda_replicas(2, spec["L"])
    rng = np.random.Generator(np.random.PCG64(W.SEED_BASE + 91))
    small = W.random_polyhex(rng, 20)
    ang, xy, h = W.uniform_batch_poses(1, L=spec["L"], hmin=3.5, hmax=3.5, seed=W.SEED_BASE + 92)
    Xs = W.place_on_surface(small, ang[0], xy[0], h[0])
    cases = [(mol, X[1]), (small, Xs)]
    assert cases[1][0].natoms < cases[0][0].natoms
    for m, x in cases:
        hm = S.Molecule.build(m.elems, x, m.charges, bonds=m.bonds)
        b_ref = S.Batch(ctx, [hm], [0], x, grid=g, rdamp=RDAMP)
        b_fix = S.Batch(ctx, [hm], [0], x, grid=g, rdamp=RDAMP)
        r_ref = b_ref.relax_plain(0, fp, 3000, 1e-3)
        r_fix = _noalloc(b_fix, 0, fp, 3000, 1e-3)
        assert r_fix["status"] == S.SMR_SYS_CONVERGED
        assert r_fix.tobytes() == r_ref.tobytes(), (m.natoms, r_fix, r_ref)
        assert np.array_equal(b_fix.positions(), b_ref.positions())
        b_ref.close()
        b_fix.close()
    g.close()


def test_plain_noalloc_errors(ctx):
    import surfmol_amd as S
    fp = S.FireParams(**fire_params_internal())
    mol, X = ptcda_replicas(1, 12.0)
    hm = S.Molecule.build(mol.elems, X[0], mol.charges, bonds=mol.bonds)
    b = S.Batch(ctx, [hm], [0], X[0], grid=None, rdamp=RDAMP)
    with pytest.raises(S.SmrError, match="isys"):
        _noalloc(b, 3, fp, 10, 1e-3)
    b.close()
