"""Hand-derived pins of the UFF parameter branches that the workloads reach
(CPU): Uff::setup_params (molff/src/uff.rs:699-887) with the UFF formulas of
moltopo/src/params.rs:607-647.

Every expected value is written out as a number, derived from the columns of
the parameter tables (ElementTypes.dat: U_uff, V_uff; AtomTypes.dat: R_uff,
theta0) and the closed forms, so that the oracle's restatement
(oracle/oracle.py setup_params) and the library's C++ host set-up
(host_topology.cpp, through smh_molecule_build) are both checked against the
reference's definitions rather than against each other.

Constants:
  KCAL = 4.1840 / 60.2214076 / 1.602176634 = 0.043364104241800934 eV
  U_C = U_N = U_O = 0.08672820848 eV (= 2 kcal/mol)
  V_N = 0.01951384691, V_O = 0.00078055388 eV
A dihedral's V is divided by 0.5 (nb2 - 1)(nb3 - 1), nb = neighbour counts
of its two central atoms (uff.rs:841-846); an inversion's K by 3 (uff.rs:882).
"""
import numpy as np
import pytest

KCAL = 0.043364104241800934

# (type2, type3, nb2, nb3) -> the rows (V, d, n) that branch can give
DIH = {
    # sp2-sp2 (uff.rs:835-838): V = 5 sqrt(U2 U3)(1 + 4.18 ln 1) = 5 U = 0.4336410424
    ("C_R", "C_R", 3, 3): {(0.2168205212, -1.0, 2.0)},           # / 2
    ("C_R", "N_R", 3, 2): {(0.4336410424, -1.0, 2.0)},           # / 1
    ("N_R", "N_R", 2, 2): {(0.8672820848, -1.0, 2.0)},           # / 0.5
    # sp3-sp2, sp3 not group 16 (uff.rs:805-833): (KCAL, -1, 6); when the sp2
    # atom has another sp2 neighbour: (2 KCAL, +1, 3)
    ("C_R", "N_3", 3, 3): {(0.021682052120900467, -1.0, 6.0),    # KCAL / 2
                           (0.043364104241800934, 1.0, 3.0)},    # 2 KCAL / 2
    ("N_3", "N_R", 3, 2): {(0.043364104241800934, -1.0, 6.0),    # KCAL / 1
                           (0.08672820848360187, 1.0, 3.0)},     # 2 KCAL / 1
    # sp3-sp2, sp3 = O (group 16, uff.rs:807-812): (5 sqrt(U U)(1 + 4.18 ln 1), +1, 2),
    # overridden by (2 KCAL, +1, 3) when the sp2 atom has another sp2 neighbour
    ("C_R", "O_3", 3, 2): {(0.4336410424, 1.0, 2.0),             # / 1
                           (0.08672820848360187, 1.0, 3.0)},     # 2 KCAL / 1
    ("N_R", "O_3", 2, 2): {(0.8672820848, 1.0, 2.0),             # / 0.5
                           (0.17345641696720374, 1.0, 3.0)},     # 2 KCAL / 0.5
    # sp3-sp3 (uff.rs:787-803): (sqrt(V2 V3), +1, 3)
    ("N_3", "N_3", 3, 3): {(0.009756923455, 1.0, 3.0)},          # V_N / 2
    ("N_3", "O_3", 3, 2): {(0.0039027693909999997, 1.0, 3.0)},   # sqrt(V_N V_O) / 1
    # group-16 sp3-sp3: (sqrt(KCAL * 2 * 2), +1, 2) = 2 sqrt(KCAL) / 0.5
    ("O_3", "O_3", 2, 2): {(0.832961984648048, 1.0, 2.0)},
}

# inversion rows (K/3, C0, C1, C2) by central type and "has an O_2 neighbour"
INV = {
    ("C_R", False): (0.08672820848360187, 1.0, -1.0, 0.0),       # 6 KCAL / 3
    ("C_R", True): (0.7227350706966823, 1.0, -1.0, 0.0),         # carbonyl: 50 KCAL / 3
    ("N_R", False): (0.08672820848360187, 1.0, -1.0, 0.0),       # 6 KCAL / 3 (uff.rs:866-868)
    ("N_3", False): (0.0, 0.0, 0.0, 0.0),                        # sp3 N: none (uff.rs:869-872)
}

# angle coefficients (c0, c1, c2, c3) by central type:
# sp2/R (uff.rs:753): (1, 0, 0, -1) with k = kappa / 9;
# sp3 (params.rs:642-647): c2 = 1 / (4 sin^2 t0), c1 = -4 c2 cos t0,
#   c0 = c2 (2 cos^2 t0 + 1), c3 = 0 with k = kappa;
#   N_3 t0 = 106.7 deg: cos t0 = -0.28736...; O_3 t0 = 104.51 deg.
ANG_C = {
    "C_R": (1.0, 0.0, 0.0, -1.0),
    "N_R": (1.0, 0.0, 0.0, -1.0),
    "N_3": (0.31750647017248834, 0.31322544566569244, 0.2725021567241628, 0.0),
    "O_3": (0.30023455477619093, 0.26733059541642124, 0.2667448515920637, 0.0),
}
# kappa = 28.7989689090648 Qi Qk / rik^5 (3 rij rjk sin^2 t0 - rik^2 cos t0),
# rik^2 = rij^2 + rjk^2 - 2 rij rjk cos t0, with the bond lengths
# r = Ri + Rj - r_EN (bond order 1): C_R-C_R 1.458, C_R-H_ 1.0814180403,
# C_R-O_3 1.3663113864, C_R-N_3 / C_R-O_2 from the same formula.
ANG_K = {
    ("C_R", ("C_R", "C_R")): 0.9079580101369287,   # kappa / 9, Q_C = 1.912
    ("C_R", ("C_R", "H_")): 0.4975170018351037,    # Q_H = 0.712
    ("C_R", ("C_R", "O_2")): 1.230116439634041,    # Q_O = 2.300
    ("N_3", ("C_R", "C_R")): 11.990261266075363,
    ("O_3", ("C_R", "C_R")): 14.345761119283452,
}


def _nb(neighs, i):
    return int(sum(1 for x in neighs[i] if x >= 0))


def _dih_rows(types, neighs, dihedrals, dih_params):
    out = []
    for d, p in zip(dihedrals, dih_params):
        key = (types[d[1]], types[d[2]], _nb(neighs, d[1]), _nb(neighs, d[2]))
        out.append((key, tuple(float(x) for x in p)))
    return out


def _in(row, allowed):
    return any(np.allclose(row, a, rtol=1e-12, atol=0.0) for a in allowed)


def _check_tables(types, neighs, angles, ang_params, dihedrals, dih_params, inversions, inv_params, seen):
    for key, row in _dih_rows(types, neighs, dihedrals, dih_params):
        if key not in DIH:
            key = (key[1], key[0], key[3], key[2])
        assert key in DIH, f"unpinned dihedral branch {key}"
        assert _in(row, DIH[key]), (key, row)
        seen["dih"].add(key)
    for iv, p in zip(inversions, inv_params):
        c = iv[0]
        carbonyl = any(types[n] == "O_2" for n in neighs[c] if n >= 0)
        key = (types[c], carbonyl)
        assert key in INV, f"unpinned inversion branch {key}"
        assert np.allclose(p, INV[key], rtol=1e-12, atol=0.0), (key, p)
        seen["inv"].add(key)
    for a, p in zip(angles, ang_params):
        tj = types[a[1]]
        assert tj in ANG_C, f"unpinned angle centre {tj}"
        assert np.allclose(p[1:], ANG_C[tj], rtol=1e-12, atol=1e-15), (tj, p)
        key = (tj, tuple(sorted((types[a[0]], types[a[2]]))))
        if key in ANG_K:
            assert p[0] == pytest.approx(ANG_K[key], rel=1e-12), key
            seen["ang"].add(key)


def _host_tables(S, m):
    hm = S.Molecule.build(m.elems, m.apos, m.charges, bonds=m.bonds)
    t = hm.tables()
    return t["types"], t["neighs"], t["ang_atoms"], t["ang_params"], t["dih_atoms"], t["dih_params"], \
        t["inv_atoms"], t["inv_params"]


def test_ptcda_branches(oracle, params):
    """The 38-atom C24H8O6 molecule of the uniform batch: 8 dihedrals about
    C_R-O_3 bonds take "sp3 bonded to another sp2" (uff.rs:814-826):
    (2 KCAL / (0.5 * 2 * 1), +1, 3) = (0.086728 eV, +1, 3); the 112 others are
    C_R-C_R sp2-sp2; 12 inversions sit on the four carbonyl carbons."""
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    m = W.ptcda_like()
    o = oracle.system_from_geometry(m.elems, m.apos, params, charges=m.charges, bonds=m.bonds)
    for tabs in [(o.types, o.neighs, o.angles, o.ang_params, o.dihedrals, o.dih_params, o.inversions, o.inv_params),
                 _host_tables(S, m)]:
        types, neighs, dihedrals, dp = tabs[0], tabs[1], tabs[4], tabs[5]
        rows = _dih_rows(types, neighs, dihedrals, dp)
        co = [r for k, r in rows if {k[0], k[1]} == {"C_R", "O_3"}]
        assert len(co) == 8
        for r in co:
            assert r == pytest.approx((0.08672820848360187, 1.0, 3.0), rel=1e-12)
        cc = [r for k, r in rows if k[:2] == ("C_R", "C_R")]
        assert len(cc) == 112
        carb = [p for iv, p in zip(tabs[6], tabs[7]) if any(types[n] == "O_2" for n in neighs[iv[0]] if n >= 0)]
        assert len(carb) == 12
        seen = {"dih": set(), "inv": set(), "ang": set()}
        _check_tables(*tabs, seen)
        assert ("O_3", ("C_R", "C_R")) in seen["ang"] and ("C_R", ("C_R", "O_2")) in seen["ang"]


def _tiny(oracle, params, types, bonds):
    """setup_params on a hand-made topology with the given UFF types."""
    n = len(types)
    angles = oracle.build_angles_from_bonds(n, bonds)
    dih = oracle.build_dihedrals_from_bonds(n, bonds)
    inv = oracle.build_inversions_from_bonds(n, bonds)
    neighs, _ = oracle.make_neigh_bs(n, bonds)
    nl = neighs.tolist()
    bp, ap, dp, ip = oracle.setup_params(params, types, nl, bonds, angles, dih, inv)
    return nl, dih, dp, inv, ip


@pytest.mark.parametrize("case", [
    # the sp2 atom (0) has another sp2 neighbour (2): (2 KCAL, +1, 3) / 2
    (["C_R", "N_3", "C_R", "H_", "C_R", "C_R"], [(0, 1), (0, 2), (0, 3), (1, 4), (1, 5)], (0, 1),
     (0.043364104241800934, 1.0, 3.0)),
    # the sp2 atom's other neighbours are H_ and N_3 (not sp2): (KCAL, -1, 6) / 2
    (["C_R", "N_3", "N_3", "H_", "C_R", "C_R"], [(0, 1), (0, 2), (0, 3), (1, 4), (1, 5)], (0, 1),
     (0.021682052120900467, -1.0, 6.0)),
    # ring O next to a C_R with only H_ besides: group-16 sp3-sp2 (5 U, +1, 2) / 1
    (["C_R", "O_3", "H_", "H_", "C_R"], [(0, 1), (0, 2), (0, 3), (1, 4)], (0, 1), (0.4336410424, 1.0, 2.0)),
    # ... and with a C_R neighbour besides: (2 KCAL, +1, 3) / 1
    (["C_R", "O_3", "C_R", "H_", "C_R"], [(0, 1), (0, 2), (0, 3), (1, 4)], (0, 1), (0.08672820848360187, 1.0, 3.0)),
    # pyridinic N_R next to ring O, other neighbour H_: (5 U, +1, 2) / 0.5
    (["N_R", "O_3", "H_", "C_R"], [(0, 1), (0, 2), (1, 3)], (0, 1), (0.8672820848, 1.0, 2.0)),
    # N_3-N_3: sqrt(V_N V_N) / 2
    (["N_3", "N_3", "C_R", "C_R", "C_R", "C_R"], [(0, 1), (0, 2), (0, 3), (1, 4), (1, 5)], (0, 1),
     (0.009756923455, 1.0, 3.0)),
    # N_3-O_3: sqrt(V_N V_O) / 1
    (["N_3", "O_3", "C_R", "C_R", "C_R"], [(0, 1), (0, 2), (0, 3), (1, 4)], (0, 1), (0.0039027693909999997, 1.0, 3.0)),
    # O_3-O_3: sqrt(4 KCAL) / 0.5
    (["O_3", "O_3", "C_R", "C_R"], [(0, 1), (0, 2), (1, 3)], (0, 1), (0.832961984648048, 1.0, 2.0)),
])
def test_dihedral_branch_selection(oracle, params, case):
    types, bonds, centre, expect = case
    nl, dih, dp, inv, ip = _tiny(oracle, params, types, bonds)
    rows = [tuple(p) for d, p in zip(dih, dp) if {d[1], d[2]} == set(centre)]
    assert rows
    for r in rows:
        assert r == pytest.approx(expect, rel=1e-12), (types, r)


def test_inversion_branches(oracle, params):
    """N_3 centres get no inversion (uff.rs:869-872); a 3-connected N_R (the
    nitro-like N of assign_uff.rs with two O neighbours) gets 6 kcal/mol / 3."""
    nl, dih, dp, inv, ip = _tiny(oracle, params, ["N_3", "C_R", "C_R", "C_R"], [(0, 1), (0, 2), (0, 3)])
    rows = [tuple(p) for iv, p in zip(inv, ip) if iv[0] == 0]
    assert len(rows) == 3 and all(r == (0.0, 0.0, 0.0, 0.0) for r in rows)
    nl, dih, dp, inv, ip = _tiny(oracle, params, ["N_R", "O_R", "O_R", "C_R"], [(0, 1), (0, 2), (0, 3)])
    rows = [tuple(p) for iv, p in zip(inv, ip) if iv[0] == 0]
    assert len(rows) == 3
    for r in rows:
        assert r == pytest.approx(INV[("N_R", False)], rel=1e-12)


def test_fused_ring_workload_branches(oracle, params):
    """Every dihedral, inversion and angle row of the uneven-batch molecules
    (random fused 6-ring C/N/O/H graphs, sizes log-uniform in [20, 400]) and
    of the serving stream is one of the pinned rows above, in the oracle and
    in the library's host set-up, and the workloads reach the N_3, O_3 and
    N_R branches."""
    import surfmol_amd as S
    from surfmol_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(W.SEED_BASE + 2))
    seen = {"dih": set(), "inv": set(), "ang": set()}
    for k in range(40):
        n = int(np.exp(rng.uniform(np.log(20.0), np.log(400.0))))
        m = W.random_polyhex(rng, n)
        o = oracle.system_from_geometry(m.elems, m.apos, params, charges=m.charges, bonds=m.bonds)
        _check_tables(o.types, o.neighs, o.angles, o.ang_params, o.dihedrals, o.dih_params, o.inversions,
                      o.inv_params, seen)
        _check_tables(*_host_tables(S, m), seen)
    for key in [("C_R", "N_3", 3, 3), ("N_3", "N_R", 3, 2), ("C_R", "O_3", 3, 2), ("N_3", "N_3", 3, 3),
                ("C_R", "N_R", 3, 2)]:
        assert key in seen["dih"] or (key[1], key[0], key[3], key[2]) in seen["dih"], key
    assert ("N_3", False) in seen["inv"] and ("C_R", True) in seen["inv"]
    assert ("N_3", ("C_R", "C_R")) in seen["ang"] and ("O_3", ("C_R", "C_R")) in seen["ang"]
