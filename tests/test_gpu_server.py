"""Serving loop on the GPU: continuous batching of relax requests.

A request's relaxation is independent of what else is resident: its size
class fixes its launch shape and FIRE state persists exactly across chunk
boundaries, so the result is bit-identical alone, in a mix that arrives
while others run, under memory-limited admission, and to smr_relax_batched
on a one-system batch.  Energies are FP64 at the returned positions.
"""
import time

import numpy as np
import pytest

from surface_setup import ALPHA, RDAMP, fire_params_internal, oracle_grid, oracle_system, small_grid_spec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def surf(ctx, oracle, params):
    import surfmol_amd as S
    spec = small_grid_spec()
    E, C, (el, pos, q, reqs) = oracle_grid(oracle, params, spec)
    gs = S.GridSpec(spec["nx"], spec["ny"], spec["nz"], spec["p0"], spec["step"], ALPHA, RDAMP)
    g = S.GridFF.build(ctx, gs, pos, reqs, npbc=2)
    return dict(spec=spec, grid=g, C=g.coefficients().astype(np.float64))


def _requests(n, seed, nmin=20, nmax=200):
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), n)).round().astype(int)
    mols = [W.random_polyhex(rng, int(t)) for t in sizes]
    ang, xy, h = W.uniform_batch_poses(n, L=12.0, hmin=3.0, hmax=4.5, seed=seed + 1)
    X = [W.place_on_surface(m, ang[s], xy[s], h[s]) for s, m in enumerate(mols)]
    hm = [S.Molecule.build(m.elems, X[s], m.charges, bonds=m.bonds) for s, m in enumerate(mols)]
    return mols, hm, X


def _serve(ctx, surf, hm, X, order=None, delays=None, **kw):
    import surfmol_amd as S
    fp = S.FireParams(**fire_params_internal())
    srv = S.Server(ctx, surf["grid"], fp, rdamp=RDAMP, **kw)
    order = list(range(len(hm))) if order is None else order
    ids = {}
    for q, s in enumerate(order):
        ids[s] = srv.submit(hm[s], X[s])
        if delays is not None:
            time.sleep(delays[q])
    srv.drain()
    out = {s: srv.poll(ids[s]) for s in order}
    stats = srv.stats()
    srv.close()
    return out, stats


def test_server_matches_batch_and_fp64(ctx, surf, oracle, params):
    import surfmol_amd as S
    mols, hm, X = _requests(6, 0x5EED2001)
    out, stats = _serve(ctx, surf, hm, X)
    assert stats["retired"] == 6 and stats["resident"] == 0
    fp = S.FireParams(**fire_params_internal())
    for s in range(6):
        r = out[s]
        assert r["status"] == S.SMR_SYS_CONVERGED, r
        # the same relaxation as a one-system batch, bit for bit
        b = S.Batch(ctx, [hm[s]], [0], X[s], grid=surf["grid"], rdamp=RDAMP)
        rb = b.relax_batched(fp, 2000, 1e-3)[0]
        assert rb["steps"] == r["steps"] and rb["energy"] == r["energy"], (s, rb, r)
        assert np.array_equal(b.positions(), r["positions"])
        b.close()
        # FP64 energy at the returned positions
        osys = oracle_system(oracle, params, mols[s], X[s], surf["C"], surf["spec"])
        fo, eo = oracle.eval_forces(osys, r["positions"])
        assert abs(eo.sum() - r["energy"]) / mols[s].natoms < 1e-6
        # convergence is tested on the forces before the last move (the
        # reference's order), so the returned positions carry a slightly
        # larger Fmax; the pin is that they are an FP64 minimum: FP64 FIRE
        # started there converges without moving the energy by 1e-6 eV/atom
        assert np.linalg.norm(fo, axis=1).max() < 1e-2
        rf = oracle.relax_fire(osys, oracle.FireParams(**fire_params_internal()), 3000, 1e-3, apos=r["positions"])
        assert rf["converged"]
        assert abs(rf["E"] - r["energy"]) / mols[s].natoms < 1e-6, (s, rf["E"], r["energy"])
        assert r["latency"] > 0


def test_server_same_alone_and_in_mix(ctx, surf):
    """Requests arriving while others run give the results they give alone."""
    mols, hm, X = _requests(24, 0x5EED2002)
    rng = np.random.Generator(np.random.PCG64(7))
    order = list(rng.permutation(24))
    delays = rng.exponential(0.002, 24)
    mix, st = _serve(ctx, surf, hm, X, order=order, delays=delays, chunk_steps=32)
    assert st["chunks"] > 1
    for s in [int(order[0]), int(order[5]), int(order[23])]:
        alone, _ = _serve(ctx, surf, [hm[s]], [X[s]], chunk_steps=100)
        a, m = alone[0], mix[s]
        assert a["steps"] == m["steps"] and a["energy"] == m["energy"], (s, a, m)
        assert np.array_equal(a["positions"], m["positions"])


def test_server_memory_limited_admission(ctx, surf):
    """With arenas for about 600 atoms, requests wait for memory, are all
    served, and give the same results as with ample memory."""
    mols, hm, X = _requests(12, 0x5EED2003, nmin=40, nmax=160)
    free = ctx.info()["free_bytes"]
    frac = 600 * 1014 / free
    small, st = _serve(ctx, surf, hm, X, max_atoms=200, hbm_fraction=frac, chunk_steps=50)
    assert st["atom_capacity"] < sum(m.natoms for m in mols)
    assert st["peak_resident"] < 12
    big, _ = _serve(ctx, surf, hm, X, chunk_steps=50)
    for s in range(12):
        assert small[s]["steps"] == big[s]["steps"] and small[s]["energy"] == big[s]["energy"]


def test_server_rejects_bad_requests(ctx, surf):
    import surfmol_amd as S
    mols, hm, X = _requests(2, 0x5EED2004, nmin=150, nmax=160)
    fp = S.FireParams(**fire_params_internal())
    srv = S.Server(ctx, surf["grid"], fp, max_atoms=100, rdamp=RDAMP)
    with pytest.raises(S.SmrError):
        srv.submit(hm[0], X[0])
    with pytest.raises(S.SmrError):
        srv.poll(12345)
    srv.close()
