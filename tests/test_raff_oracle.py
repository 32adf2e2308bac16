"""The multi-replica rigid-atom port relaxation on the CPU: the oracle's
restatement of the reference's packing (crates/libs/oclff/src/pack.rs) and
of one RRsp3Multi step (opencl/RRsp3.cl:1774-2047), and the library's C++
host packing (smh_raff_build) against it.

Pins:
  * the reference's own unit tests of pack.rs (pack.rs:359-428: two waters,
    neighbours, exclusions, back slots, ports) -- for the oracle and the host;
  * the packed reference geometry is a fixed point (every port tip sits on its
    neighbour, to the float32 rounding of the arms), and so is any rigid
    motion of it with the matching orientation -- a wrong quaternion
    convention or port index fails here;
  * a stretched water bond, derived by hand: the XPBD impulse
    0.1 / (w_O + w_H + 1/(K dt^2)) moves O and H along the bond in inverse
    proportion to their masses;
  * linear momentum: sum_i m_i dx_i = 0 for port and collision corrections;
  * relaxation restores the bond lengths of the reference geometry.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _water():
    """make_h2o_geometry (pack.rs:319-329) in float32."""
    ang = np.float32(np.radians(np.float32(104.5)))
    r = np.float32(0.96)
    h1 = [r * np.cos(ang * np.float32(0.5)), r * np.sin(ang * np.float32(0.5)), 0.0]
    h2 = [r * np.cos(ang * np.float32(0.5)), -r * np.sin(ang * np.float32(0.5)), 0.0]
    return ["O", "H", "H"], np.array([[0, 0, 0], h1, h2], dtype=np.float32), [(0, 1), (0, 2)]


def _xylitol():
    """The reference benchmark's molecule: bonds where d < r_i + r_j + 0.4
    with its covalent radii (raff_ocl_benchmark.rs:24-47)."""
    from oracle import oracle as O
    el, x, _ = O.read_xyz(os.path.join(GOLDEN, "xylitol.xyz"))
    rc = {"H": 0.31, "C": 0.76, "O": 0.66}
    bonds = []
    for i in range(len(el)):
        for j in range(i + 1, len(el)):
            cut = rc[el[i]] + rc[el[j]] + 0.4
            if ((x[j] - x[i]) ** 2).sum() < cut * cut:
                bonds.append((i, j))
    return el, x.astype(np.float32), bonds


@pytest.mark.parametrize("impl", ["oracle", "host"])
def test_reference_pack_unit_tests(oracle, impl):
    """pack.rs:359-428 on one water (group 64)."""
    el, x, bonds = _water()
    if impl == "oracle":
        t = oracle.raff_pack(el, x, bonds, nnode=1, group_size=64)
        elems, neighs, excl1, excl2, bk, port, kflat = (t["elems"], t["neighs"], t["excl1"], t["excl2"],
                                                        t["bk_slots"], t["port_local"], t["kflat"])
        pad = [e == "X" for e in elems]
    else:
        from surfmol_amd.raff import RaffMolecule
        m = RaffMolecule.build(el, x, bonds, nnode=1, group_size=64)
        elems, neighs, excl1, excl2, bk, port, kflat = (m.elems, m.neighs, m.excl1, m.excl2, m.bk_slots,
                                                        m.port_local[:, :, :3], m.kflat)
        pad = list(m.is_padding)
    # test_pack_two_water: O node first, then the H caps, then padding
    assert elems[:3] == ["O", "H", "H"] and pad[3] and len(elems) == 64
    # test_build_neighs_water
    assert list(neighs[0]) == [1, 2, -1, -1] and list(neighs[1]) == [0, -1, -1, -1] and neighs[2][0] == 0
    # test_exclusions_water
    assert list(excl1[0][:2]) == [1, 2] and excl2[0][0] == -1
    assert excl1[1][0] == 0 and excl2[1][0] == 2
    # test_bk_slots_water_packed: H1 <- node 0 port 0, H2 <- node 0 port 1
    assert bk[1][0] == 0 and bk[2][0] == 1
    # test_ports_from_neighs_water
    assert abs(port[0][0][0] - x[1][0]) < 1e-5 and abs(port[0][0][1] - x[1][1]) < 1e-5
    assert kflat[0][0] == 200.0


@pytest.mark.parametrize("case", ["water", "xylitol32", "xylitol_nopad"])
def test_host_pack_equals_oracle(oracle, case):
    from surfmol_amd.raff import RaffMolecule
    el, x, bonds = _water() if case == "water" else _xylitol()
    gs = {"water": 16, "xylitol32": 32, "xylitol_nopad": 0}[case]
    t = oracle.raff_pack(el, x, bonds, group_size=gs)
    m = RaffMolecule.build(el, x, bonds, group_size=gs)
    assert m.elems == t["elems"] and m.nnode == t["nnode"] and m.natoms == t["natoms"]
    for k in ("neighs", "excl1", "excl2", "bk_slots"):
        np.testing.assert_array_equal(getattr(m, k), t[k], err_msg=k)
    np.testing.assert_array_equal(m.pos3, t["pos3"])
    np.testing.assert_array_equal(m.inv_mass, t["inv_mass"])
    np.testing.assert_array_equal(m.port_local[:, :, :3], t["port_local"])
    np.testing.assert_array_equal(m.kflat, t["kflat"])
    np.testing.assert_array_equal(m.perm, t["perm"])


def test_pack_errors():
    from surfmol_amd import SmrError
    from surfmol_amd.raff import RaffMolecule
    el, x, bonds = _water()
    with pytest.raises(SmrError, match="degree"):
        RaffMolecule.build(el, x, bonds, nnode=2)
    with pytest.raises(SmrError, match="group_size"):
        RaffMolecule.build(el, x, bonds, group_size=2)
    with pytest.raises(SmrError, match="out of range"):
        RaffMolecule.build(el, x, [(0, 5)])


def _replica(oracle, el, x, bonds, gs=0, radius=None):
    t = oracle.raff_pack(el, x, bonds, group_size=gs)
    return t, oracle.RaffReplica(t, radius=radius)


def test_reference_geometry_is_a_fixed_point(oracle):
    el, x, bonds = _xylitol()
    t, rep = _replica(oracle, el, x, bonds, gs=32)
    r = rep.run(t["pos3"], nsteps=3)
    # the float32 port arms miss the FP64 differences by their rounding only
    assert r["metric"] < 1e-8
    assert np.abs(r["pos"] - t["pos3"].astype(np.float64))[:t["nreal"]].max() < 1e-7
    # a rigid motion of the whole molecule with the matching orientations
    a = 0.7
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    qz = np.array([0.0, 0.0, np.sin(a / 2), np.cos(a / 2)])
    x2 = t["pos3"].astype(np.float64) @ Rz.T + np.array([1.0, -2.0, 0.5])
    quat = np.tile(qz, (t["natoms"], 1))
    r2 = rep.run(x2, quat=quat, nsteps=2)
    assert r2["metric"] < 1e-8
    assert np.abs(r2["pos"][:t["nreal"]] - x2[:t["nreal"]]).max() < 1e-7


def test_stretched_water_bond_closed_form(oracle):
    """One step, relaxation 0.5, K = 200, dt = 0.1: H1 moved 0.1 A out along
    its bond.  w_O = 1/15.999 = 0.06250391155481339, w_H = 1/1.008 =
    0.9920634627342224 (float32), alpha = 1/(K dt^2) = 0.499999999999995:
    impulse = 0.1 / (w_O + w_H + alpha + 1e-12) = 0.0643265783483137,
    O moves +0.5 * 0.0040206627637067735 along the bond, H1
    0.5 * -0.06381604806207235; no rotation (the arm is along the bond), H2
    untouched."""
    el, x, bonds = _water()
    t, rep = _replica(oracle, el, x, bonds)
    p = t["pos3"].astype(np.float64)
    u = p[1] / np.linalg.norm(p[1])
    p[1] += 0.1 * u
    r = rep.run(p, nsteps=1)
    assert np.allclose(r["dnode"][0], 0.0040206627637067735 * u, atol=1e-15)
    assert np.allclose(r["dneigh"][0], -0.06381604806207235 * u, atol=1e-15)
    assert np.allclose(r["dneigh"][1], 0.0, atol=0)
    assert np.allclose(r["pos"][0], 0.5 * 0.0040206627637067735 * u, atol=1e-15)
    assert np.allclose(r["pos"][1], p[1] - 0.5 * 0.06381604806207235 * u, atol=1e-15)
    assert np.array_equal(r["pos"][2], p[2])
    assert np.array_equal(r["quat"][0], [0.0, 0.0, 0.0, 1.0])
    assert r["metric"] == pytest.approx(np.abs(0.0040206627637067735 * u).max(), rel=1e-14)


def _methanol():
    """make_ch3oh_geometry (pack.rs:332-348) in float32."""
    f = np.float32
    tet, oh = np.radians(f(109.5)), np.radians(f(108.5))
    ct, st = np.cos(tet), np.sin(tet)
    c120, s120 = np.cos(np.radians(f(120.0))), np.sin(np.radians(f(120.0)))
    o = [1.43, 0.0, 0.0]
    pos = [[0, 0, 0], o, [-1.09, 0, 0], [1.09 * ct, 1.09 * st, 0], [1.09 * ct, 1.09 * st * c120, 1.09 * st * s120],
           [o[0] + 0.96 * np.cos(np.pi - oh), o[1] + 0.96 * np.sin(np.pi - oh), 0.0]]
    return ["C", "O", "H", "H", "H", "H"], np.array(pos, dtype=np.float32), [(0, 1), (0, 2), (0, 3), (0, 4), (1, 5)]


@pytest.mark.parametrize("case", ["xylitol_ports", "methanol_collisions"])
def test_corrections_conserve_linear_momentum(oracle, case):
    """sum_i m_i dx_i = 0: each port impulse moves its node by P w_i and its
    neighbour by -P w_j, each collision pair by +-n dl w.  (The reference's
    exclusion lists hold at most four second neighbours; where that cuts a
    list, a pair can be excluded on one side only, so the collision case uses
    methanol, whose lists are complete.)"""
    el, x, bonds = _xylitol() if case == "xylitol_ports" else _methanol()
    t = oracle.raff_pack(el, x, bonds, group_size=32)
    rad = np.where(np.array(t["elems"]) == "X", 0.0, 0.0 if case == "xylitol_ports" else 1.4)
    rep = oracle.RaffReplica(t, radius=rad)
    rng = np.random.default_rng(5)
    p = t["pos3"].astype(np.float64) + rng.uniform(-0.2, 0.2, (32, 3))
    r = rep.run(p, nsteps=1, relaxation=1.0)
    m = np.where(t["inv_mass"] > 0, 1.0 / np.maximum(t["inv_mass"].astype(np.float64), 1e-30), 0.0)
    dx = r["pos"] - p
    dx[t["nreal"]:] = 0.0
    P = (m[:, None] * dx).sum(axis=0)
    assert np.abs(P).max() < 1e-12 * np.abs(m[:, None] * dx).sum()
    if case == "methanol_collisions":
        assert np.abs(r["dcoll"]).max() > 0       # collisions act


def test_relaxation_restores_bond_lengths(oracle):
    el, x, bonds = _xylitol()
    t, rep = _replica(oracle, el, x, bonds, gs=32)
    nb = [(i, j) for i in range(t["nreal"]) for j in t["neighs"][i] if j > i]
    p0 = t["pos3"].astype(np.float64)
    d0 = np.array([np.linalg.norm(p0[i] - p0[j]) for i, j in nb])
    rng = np.random.default_rng(9)
    p = p0 + rng.uniform(-0.15, 0.15, p0.shape)
    p[t["nreal"]:] = 0.0
    err0 = np.abs(np.array([np.linalg.norm(p[i] - p[j]) for i, j in nb]) - d0).max()
    r = rep.run(p, nsteps=3000, dt=0.1, relaxation=0.5)
    err = np.abs(np.array([np.linalg.norm(r["pos"][i] - r["pos"][j]) for i, j in nb]) - d0).max()
    assert err < 0.05 * err0, (err0, err)
