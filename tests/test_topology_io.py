"""Topology ingestion from the reference's on-disk formats (CPU).

* UFFTOPO binary, the layout buildff writes (crates/apps/buildff/src/
  main.rs:10-30 header, :125-182 writer).  An independent Python
  restatement of that layout (``_write_ufftopo`` / ``_read_ufftopo`` below,
  struct-based) writes files the library reads, and reads the files the
  library writes: a mistake in either side's field order, width or padding
  fails one direction.
* TopologyData JSON (moltopo/src/export.rs:12-26), including the CH4 example
  of crates/apps/README.md:72-88 completed to a full tetrahedron.
* A read topology equals the one built from the same geometry: same
  parameters per term, whatever order the file lists the terms in (the
  reference's dihedral order comes from a HashSet, topology.rs:141-158).
* Bad files are errors with a message, never an abort.
"""
import json
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN

MAGIC = b"UFFTOPO"
HDR = struct.Struct("<7sB7i")   # magic, version, natoms nbonds nangles ndih ninv ntypes flags


def _write_ufftopo(path, apos, types, bonds, angles, dihedrals, inversions, version=1):
    table = sorted(set(types))
    idx = [table.index(t) for t in types]
    b = bytearray(HDR.pack(MAGIC, version, len(types), len(bonds), len(angles), len(dihedrals), len(inversions),
                           len(table), 0))
    b += struct.pack(f"<{3 * len(types)}d", *np.asarray(apos, dtype=float).ravel())
    b += struct.pack(f"<{len(idx)}i", *idx)
    for rows in (bonds, angles, dihedrals, inversions):
        flat = [int(v) for r in rows for v in r]
        b += struct.pack(f"<{len(flat)}i", *flat)
    for t in table:
        b += t.encode().ljust(8, b" ")
    with open(path, "wb") as f:
        f.write(bytes(b))


def _read_ufftopo(path):
    data = open(path, "rb").read()
    magic, ver, n, nb, na, nd, ni, nt, flags = HDR.unpack_from(data, 0)
    o = HDR.size
    apos = np.array(struct.unpack_from(f"<{3 * n}d", data, o)).reshape(n, 3); o += 24 * n
    ti = struct.unpack_from(f"<{n}i", data, o); o += 4 * n
    out = {}
    for name, k, cnt in (("bonds", 2, nb), ("angles", 3, na), ("dihedrals", 4, nd), ("inversions", 4, ni)):
        out[name] = np.array(struct.unpack_from(f"<{k * cnt}i", data, o), dtype=np.int32).reshape(-1, k)
        o += 4 * k * cnt
    table = [data[o + 8 * q:o + 8 * q + 8].decode().rstrip(" ") for q in range(nt)]
    o += 8 * nt
    assert o == len(data)
    return dict(magic=magic, version=ver, flags=flags, apos=apos, types=[table[i] for i in ti], table=table, **out)


def _xyz(name):
    from oracle import oracle as O
    el, apos, _ = O.read_xyz(os.path.join(GOLDEN, name))
    return el, np.asarray(apos, dtype=float)


@pytest.fixture(scope="module")
def pentacene():
    import surfmol_amd as S
    el, X = _xyz("pentacene.xyz")
    return S.Molecule.build(el, X)


def _by_term(tab, atoms_key, par_key, ngs_key):
    return {tuple(a): (tuple(p), tuple(g)) for a, p, g in zip(tab[atoms_key], tab[par_key], tab[ngs_key])}


def _same_topology(m1, m2):
    t1, t2 = m1.tables(), m2.tables()
    for k in ("bonds", "neighs", "neigh_bs", "bon_params", "reqs"):
        np.testing.assert_array_equal(t1[k], t2[k], err_msg=k)
    assert t1["types"] == t2["types"]
    for a, p, g in (("ang_atoms", "ang_params", "ang_ngs"), ("dih_atoms", "dih_params", "dih_ngs"),
                    ("inv_atoms", "inv_params", "inv_ngs")):
        assert len(t1[a]) == len(t2[a]), a
        assert _by_term(t1, a, p, g) == _by_term(t2, a, p, g), a


def test_ufftopo_written_by_library_has_buildff_layout(pentacene, tmp_path):
    path = tmp_path / "pentacene.ufftopo"
    pentacene.write_ufftopo(path)
    f = _read_ufftopo(path)
    t = pentacene.tables()
    n, nb, na, nd, ni = pentacene.counts
    assert f["magic"] == MAGIC and f["version"] == 1 and f["flags"] == 0
    assert f["table"] == sorted(set(t["types"]))   # buildff: sorted, deduplicated type names
    assert f["types"] == t["types"]
    np.testing.assert_array_equal(f["apos"], pentacene.positions())
    np.testing.assert_array_equal(f["bonds"], t["bonds"])
    np.testing.assert_array_equal(f["angles"], t["ang_atoms"])
    np.testing.assert_array_equal(f["dihedrals"], t["dih_atoms"])
    np.testing.assert_array_equal(f["inversions"], t["inv_atoms"])
    assert (len(f["bonds"]), len(f["angles"]), len(f["dihedrals"]), len(f["inversions"])) == (nb, na, nd, ni)


def test_ufftopo_written_independently_reads_back(pentacene, tmp_path):
    import surfmol_amd as S
    t = pentacene.tables()
    rng = np.random.default_rng(7)
    # the terms in another order, as the reference's HashSet would list them
    dih = t["dih_atoms"][rng.permutation(len(t["dih_atoms"]))]
    ang = t["ang_atoms"][rng.permutation(len(t["ang_atoms"]))]
    path = tmp_path / "shuffled.ufftopo"
    _write_ufftopo(path, pentacene.positions(), t["types"], t["bonds"], ang, dih, t["inv_atoms"])
    m = S.Molecule.read_ufftopo(path)
    _same_topology(m, pentacene)
    np.testing.assert_array_equal(m.tables()["dih_atoms"], dih)   # the file's order is kept
    np.testing.assert_array_equal(m.positions(), pentacene.positions())


def test_ufftopo_round_trip_is_exact(pentacene, tmp_path):
    import surfmol_amd as S
    path = tmp_path / "rt.ufftopo"
    pentacene.write_ufftopo(path)
    m = S.Molecule.read_ufftopo(path)
    for k, v in pentacene.tables().items():
        if k == "types":
            assert m.types == v
        else:
            np.testing.assert_array_equal(m.tables()[k], v, err_msg=k)


def test_json_round_trip_and_layout(pentacene, tmp_path):
    import surfmol_amd as S
    path = tmp_path / "pentacene.json"
    pentacene.write_json(path)
    d = json.load(open(path))
    assert set(d) == {"natoms", "elements", "positions", "bonds", "angles", "dihedrals", "inversions", "bond_params",
                      "angle_params", "dihedral_params", "inversion_params", "atom_params"}
    assert d["natoms"] == pentacene.natoms
    np.testing.assert_array_equal(np.array(d["positions"]), pentacene.positions())
    np.testing.assert_array_equal(np.array(d["bonds"]), pentacene.tables()["bonds"])
    m = S.Molecule.read_json(path)
    for k, v in pentacene.tables().items():
        if k == "types":
            assert m.types == v
        else:
            np.testing.assert_array_equal(m.tables()[k], v, err_msg=k)


def test_json_readme_example_methane(tmp_path):
    """CH4 as in crates/apps/README.md:72-88 (the positions there are
    (0,0,0) and (0.629,0.629,0.629), ...: a tetrahedron)."""
    import surfmol_amd as S
    a = 0.629
    pos = [[0, 0, 0], [a, a, a], [-a, -a, a], [-a, a, -a], [a, -a, -a]]
    d = {"natoms": 5, "elements": ["C", "H", "H", "H", "H"], "positions": pos,
         "bonds": [[0, 1], [0, 2], [0, 3], [0, 4]],
         "angles": [[1, 0, 2], [1, 0, 3], [1, 0, 4], [2, 0, 3], [2, 0, 4], [3, 0, 4]],
         "dihedrals": [], "inversions": [], "bond_params": [], "angle_params": [], "dihedral_params": [],
         "inversion_params": [], "atom_params": []}
    path = tmp_path / "ch4.json"
    path.write_text(json.dumps(d, indent=2))
    m = S.Molecule.read_json(path)
    assert m.counts == (5, 4, 6, 0, 0)
    assert m.types == ["C_3", "H_", "H_", "H_", "H_"]
    ref = S.Molecule.build(d["elements"], np.array(pos, dtype=float), bonds=d["bonds"])
    _same_topology(m, ref)
    # sp3 carbon: tetrahedral angle term, no inversion
    tab = m.tables()
    assert np.all(tab["ang_params"][:, 0] > 0)


def test_json_without_terms_is_bonds_only(tmp_path):
    import surfmol_amd as S
    d = {"natoms": 2, "elements": ["H", "H"], "positions": [[0, 0, 0], [0.74, 0, 0]], "bonds": [[0, 1]]}
    path = tmp_path / "h2.json"
    path.write_text(json.dumps(d))
    m = S.Molecule.read_json(path, charges=[0.1, -0.1])
    assert m.counts == (2, 1, 0, 0, 0)
    np.testing.assert_allclose(m.tables()["reqs"][:, 2], [0.1, -0.1])


@pytest.mark.parametrize("case", ["magic", "version", "truncated", "extra", "index", "not_bonded", "type"])
def test_bad_ufftopo_is_an_error(pentacene, tmp_path, case):
    import surfmol_amd as S
    t = pentacene.tables()
    path = tmp_path / f"{case}.ufftopo"
    types, dih = list(t["types"]), t["dih_atoms"].copy()
    if case == "index":
        dih[0, 3] = 999
    if case == "not_bonded":
        dih[0] = [dih[0, 0], dih[0, 1], dih[0, 3], dih[0, 2]]   # i-j-l-k: l-k bonded but j-l not
    if case == "type":
        types[0] = "Xx_9"
    _write_ufftopo(path, pentacene.positions(), types, t["bonds"], t["ang_atoms"], dih, t["inv_atoms"],
                   version=2 if case == "version" else 1)
    data = bytearray(path.read_bytes())
    if case == "magic":
        data[0:7] = b"UFFTOPX"
    if case == "truncated":
        data = data[:-5]
    if case == "extra":
        data += b"\0\0\0\0"
    path.write_bytes(bytes(data))
    with pytest.raises(S.SmrError):
        S.Molecule.read_ufftopo(path)


@pytest.mark.parametrize("text", ['{"natoms": 2, "elements": ["H"], "positions": [[0,0,0],[1,0,0]]}',
                                  '{"natoms": 2, "elements": ["H","H"], "positions": [[0,0,0],[1,0,0]], '
                                  '"bonds": [[0,1,2]]}',
                                  '{"natoms": 2, "elements": ["H","H"], "positions": [[0,0,0],[1,0,0]], '
                                  '"bonds": [[0,5]]}',
                                  '{"natoms": 1, "elements": ["C"], "positions": [[0,0,0]]',
                                  '[1, 2, 3]'])
def test_bad_json_is_an_error(tmp_path, text):
    import surfmol_amd as S
    path = tmp_path / "bad.json"
    path.write_text(text)
    with pytest.raises(S.SmrError):
        S.Molecule.read_json(path)


def test_missing_file_is_an_error(tmp_path):
    import surfmol_amd as S
    with pytest.raises(S.SmrError):
        S.Molecule.read_ufftopo(tmp_path / "nope.ufftopo")
    with pytest.raises(S.SmrError):
        S.Molecule.read_json(tmp_path / "nope.json")
