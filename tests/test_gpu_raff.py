"""Multi-replica rigid-atom port relaxation on the GPU (include/surfmol_raff.h)
against the FP64 oracle's restatement of the reference step.

Set-up of the reference's RAFF benchmark (crates/libs/oclff/src/bin/
raff_ocl_benchmark.rs): xylitol packed into 32-atom groups (22 atoms, 10
nodes), port stiffness 200, its LCG perturbations (0.2 A, 0.3 A for H),
StepConfig dt 0.1, relaxation 0.5.  The device computes in float32 as the
reference does; positions after k steps must agree with the FP64 oracle
within POS_ATOL (float32 resolves ~1e-6 A at these coordinates), a step's
corrections within 1e-5 of their largest component.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

POS_ATOL = 5e-5


def _setup(ctx, oracle, nsys, group_size=32, radius=None, fixmask=None, mol=None):
    from surfmol_amd import workloads as W
    from surfmol_amd.raff import RaffMolecule, RaffReplicas
    el, x, bonds = mol if mol is not None else W.xylitol_raff_molecule(os.path.join(GOLDEN, "xylitol.xyz"))
    m = RaffMolecule.build(el, x, bonds, group_size=group_size)
    X = W.raff_perturbed_replicas(m.pos3, [e == "H" for e in m.elems], m.is_padding, nsys)
    rad = np.zeros(m.natoms, np.float32) if radius is None else radius(m)
    fm = None if fixmask is None else fixmask(m)
    dev = RaffReplicas.from_molecule(ctx, m, X, nsys, radius=rad, fixmask=fm)
    t = oracle.raff_pack(el, x, bonds, group_size=group_size)
    orc = oracle.RaffReplica(t, radius=rad, fixmask=fm)
    return m, X, dev, orc


@pytest.mark.parametrize("k", [1, 10, 100])
def test_steps_match_oracle(ctx, oracle, k):
    from surfmol_amd.raff import StepConfig
    m, X, dev, orc = _setup(ctx, oracle, 64)
    cfg = StepConfig(dt=0.1, k_coll=0.0, relaxation=0.5)
    dev.step(cfg, k)
    Xd = dev.positions()
    worst = 0.0
    for s in (0, 1, 31, 62, 63):
        r = orc.run(X[s], nsteps=k, dt=0.1, relaxation=0.5)
        d = np.abs(Xd[s, :m.nreal] - r["pos"][:m.nreal]).max()
        worst = max(worst, d)
        assert d < POS_ATOL, (k, s, d)
        if k == 1:
            dn = dev.dpos_node_replica(s)
            scale = np.abs(r["dnode"]).max()
            assert np.abs(dn - r["dnode"]).max() < 1e-5 * scale, (s, np.abs(dn - r["dnode"]).max(), scale)
            q = dev.quat_replica(s)
            assert np.abs(q[:m.nnode] - r["quat"][:m.nnode]).max() < 1e-6
            assert dev.metric()[s] == pytest.approx(r["metric"], rel=1e-5)
    assert np.isnan(Xd[0, m.nreal:]).all()       # padding atoms keep NaN positions
    print(f"k={k}: max position deviation {worst:.2e} A")
    dev.close()


def test_collisions_momentum_and_fixmask(ctx, oracle):
    """Sphere collisions (covalent radii + 0.5 A: 1-4 pairs of xylitol
    collide), heavy-ball momentum beta = 0.3 and a fix mask (atom 0 frozen in
    x and rotation, the last H under the z pin of bit 8)."""
    from surfmol_amd.raff import StepConfig
    rc = {"H": 0.31, "C": 0.76, "O": 0.66, "X": 0.0}

    def radius(m):
        return np.array([rc[e] + (0.5 if e != "X" else 0.0) for e in m.elems], np.float32)

    def fixmask(m):
        f = np.zeros(m.natoms, np.int32)
        f[0] = 1
        f[m.nreal - 1] = 8
        return f

    m, X, dev, orc = _setup(ctx, oracle, 16, radius=radius, fixmask=fixmask)
    cfg = StepConfig(dt=0.1, relaxation=0.5, momentum_beta=0.3)
    for k in (1, 20):
        if k > 1:
            dev.upload_state(X, m.inv_mass)
        dev.step(cfg, k)
        Xd = dev.positions()
        for s in (0, 5, 15):
            r = orc.run(X[s], nsteps=k, dt=0.1, relaxation=0.5, beta=0.3)
            assert np.abs(Xd[s, :m.nreal] - r["pos"][:m.nreal]).max() < POS_ATOL, (k, s)
            if k == 1:
                dc = dev.dpos_coll_replica(s)
                assert np.abs(r["dcoll"]).max() > 0
                assert np.abs(dc - r["dcoll"]).max() < 1e-5 * np.abs(r["dcoll"]).max()
        assert (Xd[:, 0, 0] == X[:, 0, 0]).all()            # frozen in x
    dev.close()


def test_relax_retires_each_replica(ctx, oracle):
    """Each replica stops at the step whose largest correction component
    falls below tol; the FP64 oracle stops within a few steps of it (float32
    and FP64 cross the threshold at nearly the same step) at the same
    geometry."""
    from surfmol_amd.raff import StepConfig
    m, X, dev, orc = _setup(ctx, oracle, 256)
    cfg = StepConfig(dt=0.1, relaxation=0.5)
    tol = 2e-4
    steps, status, metric = dev.relax(cfg, 4000, tol)
    assert (status == 1).all(), np.unique(status, return_counts=True)
    assert (metric < tol).all()
    assert len(np.unique(steps)) > 10                  # replicas retire at their own steps
    Xd = dev.positions()
    for s in (0, 100, 255):
        r = orc.run(X[s], nsteps=4000, dt=0.1, relaxation=0.5, tol=tol)
        assert abs(r["steps"] - steps[s]) <= max(3, 0.02 * r["steps"]), (s, r["steps"], steps[s])
        assert np.abs(Xd[s, :m.nreal] - r["pos"][:m.nreal]).max() < 1e-3
    # a second relax call starts where the first stopped: nothing left to do
    steps2, status2, _ = dev.relax(cfg, 4000, tol)
    assert (steps2 == 1).all() and (status2 == 1).all()
    dev.close()


def test_replica_alone_and_in_batch_bit_identical(ctx, oracle):
    from surfmol_amd.raff import RaffReplicas, StepConfig
    m, X, dev, orc = _setup(ctx, oracle, 64)
    cfg = StepConfig(dt=0.1, relaxation=0.5)
    dev.relax(cfg, 3000, 2e-4)
    Xb = dev.positions()
    one = RaffReplicas.from_molecule(ctx, m, X[37:38], 1)
    one.relax(cfg, 3000, 2e-4)
    assert np.array_equal(one.positions()[0], Xb[37], equal_nan=True)
    assert np.array_equal(one.quat_replica(0), dev.quat_replica(37))
    one.close()
    dev.close()


@pytest.mark.parametrize("target", [40, 90, 180])
def test_larger_molecules_match_oracle(ctx, oracle, target):
    """Fused-ring molecules with H caps in 64-, 128- and 256-lane segments
    (several waves per replica)."""
    from surfmol_amd import workloads as W
    from surfmol_amd.raff import StepConfig
    rng = np.random.Generator(np.random.PCG64(W.SEED_BASE + 300 + target))
    mol = W.random_polyhex(rng, target)
    b = [tuple(int(v) for v in e) for e in mol.bonds]
    m, X, dev, orc = _setup(ctx, oracle, 8, group_size=0, mol=(mol.elems, mol.apos.astype(np.float32), b))
    cfg = StepConfig(dt=0.1, relaxation=0.5)
    assert m.natoms == len(mol.elems) and m.natoms <= 256
    dev.step(cfg, 20)
    Xd = dev.positions()
    met = dev.metric()
    for s in (0, 7):
        r = orc.run(X[s], nsteps=20, dt=0.1, relaxation=0.5)
        assert np.abs(Xd[s] - r["pos"]).max() < POS_ATOL, (target, s)
        assert met[s] == pytest.approx(r["metric"], rel=1e-4)
    dev.close()


def test_errors(ctx):
    from surfmol_amd import SmrError
    from surfmol_amd.raff import RaffReplicas, StepConfig
    with pytest.raises(SmrError, match="natoms_per_sys"):
        RaffReplicas(ctx, 300, 4)
    r = RaffReplicas(ctx, 8, 4)
    with pytest.raises(SmrError, match="upload"):
        r.step(StepConfig(), 1)
    with pytest.raises(SmrError, match="neighs"):
        r.upload_neighs_and_exclusions(np.full((8, 4), 9), -np.ones((8, 4)), -np.ones((8, 4)))
    r.close()


def test_server_continuous_batching(ctx, oracle):
    """200 requests arrive while others relax (64 slots, 32-step chunks):
    every result equals the one-shot relaxation of the same replica bit for
    bit (positions, steps, status, metric), whatever it shared the device
    with; the server queued requests and ran many chunks."""
    import time
    from surfmol_amd.raff import RaffReplicas, RaffServer, StepConfig
    m, X, dev, orc = _setup(ctx, oracle, 200)
    cfg = StepConfig(dt=0.1, relaxation=0.5)
    tol = 2e-4
    steps, status, metric = dev.relax(cfg, 3000, tol)
    Xb = dev.positions()
    tmpl = RaffReplicas.from_molecule(ctx, m, X[:1], 1)
    srv = RaffServer(tmpl, capacity=64, chunk_steps=32, max_steps=3000, tol=tol, cfg=cfg)
    rng = np.random.default_rng(3)
    ids = []
    for s in range(200):
        ids.append(srv.submit(X[s]))
        if s % 10 == 0:
            time.sleep(float(rng.uniform(0, 0.002)))
    srv.drain()
    st = srv.stats()
    assert st["retired"] == 200 and st["resident"] == 0 and st["peak_resident"] <= 64 and st["chunks"] > 5
    for s, rid in enumerate(ids):
        r = srv.poll(rid)
        assert r is not None
        assert r["steps"] == steps[s] and r["status"] == status[s] and r["metric"] == metric[s], s
        assert np.array_equal(r["positions"][:m.nreal], Xb[s, :m.nreal]), s
        assert r["latency"] >= 0
    with pytest.raises(Exception):
        srv.poll(ids[0])                      # forgotten after the first poll
    srv.close()
    tmpl.close()
    dev.close()
