"""FP64 oracle for SurfMol's multi-system relaxation path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker (or as the timed CPU
baseline).  The product (``surfmol_amd``) never imports it.

Two independent halves:

* ``oracle.c`` (built by ``oracle/Makefile`` into ``oracle/_build/liboracle.so``):
  FP64 restatement of the force field and integrators -- UFF bonded terms
  (molff/src/uff.rs:440-637), the non-bonded pair term in both reference
  conventions (opencl/UFF.cl:57-95,1023-1204 and molff/src/nonbonded.rs:
  166-243), the tricubic GridFF sampler (opencl/gridff_eval.cl:14-286), the
  reference grid build (opencl/gridff_build.cl:908-1062), the B-spline fit,
  damped MD (surfmol/src/mol_world.rs:141-164) and FIRE
  (surfmol/tests/relax_pentacene_uff.rs:142-193, molff/src/raff.rs:1040-1062).
* this file: a pure-Python restatement of the host-side set-up -- topology
  enumeration (moltopo/src/topology.rs:101-182), neighbour tables
  (moltopo/src/molecular.rs:34-78), UFF typing (moltopo/src/assign_uff.rs:
  8-123), the .dat loaders and UFF formulas (moltopo/src/params.rs:120-657)
  and parameter assignment (molff/src/uff.rs:250-315,699-887).

How each part is pinned (tests/test_oracle_pins.py, test_uff_branch_pins.py,
test_integrator_pins.py, test_layer_oracle.py):

=====================  ======================================================
part                   pin
=====================  ======================================================
bond                   H2 known answer F=20, E_atom=1 (oclff/tests/
                       test_uff_cl.rs:8-56); l0(C_R-C_R)=1.458
                       (doc/topical_audit/uff.md:67)
angle/dihedral/inv.    finite-difference gradient of the restated energy
                       (angle: with the reference's c1 omission undone);
                       sum F = 0, sum r x F = 0; zero dihedral/inversion
                       energy on the planar pentacene (uff.md:91-93)
UFF params/typing      pentacene golden numbers E_bond=3.83, E_angle=0.01,
                       E_dih=0, E_inv=0, Fz=-11.7 (uff.md:91-93); every
                       dihedral/inversion/angle branch the workload molecules
                       reach, against values written out from the table
                       columns and closed forms (tests/test_uff_branch_pins.py)
non-bonded             closed forms at r=R0 (E=-E0, F=0), Coulomb closed
                       form, FD gradient, conventions cross-checked
GridFF sampler         partition of unity, exact reproduction of cubic
                       fields, periodic wrap, FD gradient
grid build / fit       brute force on tiny inputs vs direct Morse/Coulomb
                       sums; S*C = E after the fit
FIRE step              a 30-step 1-D harmonic-well sequence derived in exact
                       arithmetic (n_min transition, dt growth and cap, uphill
                       reset); mixing and the flim clamp on one step
                       (tests/test_integrator_pins.py)
damped MD step         one-step closed form with clamp, damping and the
                       after-the-move purge; momentum conservation
                       (relax_xylitol_uff.rs:228-229)
=====================  ======================================================

Functions with no such pin: none.  Branches of setup_params that no workload
reaches (S atoms, P inversions, sp1 centres) are restated but not pinned.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

COULOMB_CONST = 14.3996448915
KCAL_TO_EV = 4.1840 / 60.2214076 / 1.602176634
EXCL_MAX = 16
NB_OFF, NB_GPU, NB_CPU = 0, 1, 2


# ---------------------------------------------------------------------------
# build + load the C part
# ---------------------------------------------------------------------------
def build(force: bool = False) -> str:
    """Compile oracle.c with gcc via oracle/Makefile."""
    src = os.path.join(_HERE, "oracle.c")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        d, i = ctypes.c_double, ctypes.c_int
        L.orc_uff_eval.argtypes = [P(OrcSystem), P(d), P(d), P(d)]
        L.orc_nb_eval.argtypes = [P(OrcSystem), P(d), P(d)]
        L.orc_nb_eval.restype = d
        L.orc_grid_eval.argtypes = [P(OrcSystem), P(d), P(d)]
        L.orc_grid_eval.restype = d
        L.orc_eval_forces.argtypes = [P(OrcSystem), P(d), P(d), P(d)]
        L.orc_relax_fire.argtypes = [P(OrcSystem), P(d), P(d), P(OrcFire), i, d, P(d)]
        L.orc_relax_fire.restype = i
        L.orc_run_md.argtypes = [P(OrcSystem), P(d), P(d), i, d, d, d, d]
        L.orc_run_md.restype = i
        L.orc_md_steps.argtypes = [P(OrcSystem), P(d), P(d), i, d, d, d]
        L.orc_fire_step.argtypes = [i, P(d), P(d), P(d), P(d), P(OrcFire)]
        L.orc_fire_step.restype = d
        L.orc_md_step.argtypes = [i, P(d), P(d), P(d), d, d, d, P(d)]
        L.orc_make_second_neighs.argtypes = [i, P(i), P(i)]
        L.orc_make_second_neighs.restype = i
        L.orc_fe3d_pbc_comb.argtypes = [P(d), i, i, i, P(d), P(d), P(d)]
        L.orc_plqh.argtypes = [P(d), d, P(d)]
        L.orc_gridff_build.argtypes = [i, P(d), P(d), P(d), i, i, i, P(d), P(i), d, d, P(d)]
        L.orc_raff_step.argtypes = [P(OrcRaff), d, d, d, d, i, P(d), P(d), P(d), P(d), P(d), P(d), P(d), P(d)]
        L.orc_raff_step.restype = d
        L.orc_gridff_build_nodes.argtypes = [i, P(d), P(d), P(d), i, i, i, P(d), P(i), d, d, i, P(i), P(d)]
        L.orc_bspline_apply.argtypes = [i, i, i, P(d), P(d)]
        L.orc_bspline_fit.argtypes = [i, i, i, P(d), P(d)]
        _lib = L
    return _lib


class OrcSystem(ctypes.Structure):
    _fields_ = [
        ("natoms", ctypes.c_int), ("nbonds", ctypes.c_int), ("nangles", ctypes.c_int),
        ("ndih", ctypes.c_int), ("ninv", ctypes.c_int),
        ("neighs", ctypes.POINTER(ctypes.c_int)), ("neigh_bs", ctypes.POINTER(ctypes.c_int)),
        ("bon_params", ctypes.POINTER(ctypes.c_double)),
        ("ang_atoms", ctypes.POINTER(ctypes.c_int)), ("ang_ngs", ctypes.POINTER(ctypes.c_int)),
        ("ang_params", ctypes.POINTER(ctypes.c_double)),
        ("dih_atoms", ctypes.POINTER(ctypes.c_int)), ("dih_ngs", ctypes.POINTER(ctypes.c_int)),
        ("dih_params", ctypes.POINTER(ctypes.c_double)),
        ("inv_atoms", ctypes.POINTER(ctypes.c_int)), ("inv_ngs", ctypes.POINTER(ctypes.c_int)),
        ("inv_params", ctypes.POINTER(ctypes.c_double)),
        ("nb_mode", ctypes.c_int),
        ("reqs", ctypes.POINTER(ctypes.c_double)), ("excl", ctypes.POINTER(ctypes.c_int)),
        ("rdamp", ctypes.c_double), ("fmax_nb", ctypes.c_double), ("clamp_nb", ctypes.c_int),
        ("rcut2", ctypes.c_double),
        ("grid", ctypes.POINTER(ctypes.c_double)),
        ("gnx", ctypes.c_int), ("gny", ctypes.c_int), ("gnz", ctypes.c_int),
        ("gp0", ctypes.c_double * 3), ("ginv_step", ctypes.c_double * 3),
        ("alpha_morse", ctypes.c_double),
        ("mol_id", ctypes.POINTER(ctypes.c_int)), ("rcut_inter2", ctypes.c_double), ("layer_cells", ctypes.c_int),
        ("rswitch_inter", ctypes.c_double), ("pbc_lx", ctypes.c_double), ("pbc_ly", ctypes.c_double),
        ("fscale", ctypes.POINTER(ctypes.c_double)),
    ]


class OrcRaff(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("nnode", ctypes.c_int), ("radius", ctypes.POINTER(ctypes.c_double)),
        ("neighs", ctypes.POINTER(ctypes.c_int)), ("excl1", ctypes.POINTER(ctypes.c_int)),
        ("excl2", ctypes.POINTER(ctypes.c_int)), ("bk_slots", ctypes.POINTER(ctypes.c_int)),
        ("fixmask", ctypes.POINTER(ctypes.c_int)), ("port_local", ctypes.POINTER(ctypes.c_double)),
        ("kflat", ctypes.POINTER(ctypes.c_double)), ("inv_mass", ctypes.POINTER(ctypes.c_double)),
    ]


class OrcFire(ctypes.Structure):
    _fields_ = [
        ("dt0", ctypes.c_double), ("dt_max", ctypes.c_double), ("alpha0", ctypes.c_double),
        ("n_min", ctypes.c_int), ("f_inc", ctypes.c_double), ("f_dec", ctypes.c_double),
        ("f_alpha", ctypes.c_double), ("flim", ctypes.c_double),
    ]


def _pi(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return a


# ---------------------------------------------------------------------------
# .dat parameter tables  (moltopo/src/params.rs:120-291)
# ---------------------------------------------------------------------------
@dataclass
class ElementType:
    name: str
    r_vdw: float
    e_vdw: float
    q_uff: float
    u_uff: float
    v_uff: float
    e_aff: float


@dataclass
class AtomType:
    name: str
    element: str
    r_uff: float
    r_vdw: float
    e_vdw: float
    q_base: float
    h_b: float
    a_ss: float


class Params:
    """Restates Params::load_element_types/load_atom_types (params.rs:232-258)."""

    def __init__(self, data_dir: str):
        self.elements: dict[str, ElementType] = {}
        self.atom_types: dict[str, AtomType] = {}
        for line in open(os.path.join(data_dir, "ElementTypes.dat")):
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            p = line.split()
            if len(p) < 10:
                continue
            u = float(p[10]) if len(p) >= 11 else 0.0
            v = float(p[11]) if len(p) >= 12 else 0.0
            eaff = float(p[12]) if len(p) >= 16 else 0.0
            self.elements[p[0]] = ElementType(p[0], float(p[7]), float(p[8]), float(p[9]), u, v, eaff)
        for line in open(os.path.join(data_dir, "AtomTypes.dat")):
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            p = line.split()
            if len(p) < 13:
                continue
            a_ss = float(p[13]) if len(p) >= 19 else 0.0
            self.atom_types[p[0]] = AtomType(p[0], p[2], float(p[8]), float(p[9]), float(p[10]), float(p[11]),
                                             float(p[12]), a_ss)

    def element_of_atom_type(self, tname):
        at = self.atom_types.get(tname)
        if at is None:
            return None
        return self.elements.get(at.element)


def get_reqh(params: Params, tname: str):
    """params.rs:595-604: (RvdW, sqrt(EvdW), Qbase, Hb)."""
    at = params.atom_types.get(tname)
    if at is not None:
        return [at.r_vdw, math.sqrt(at.e_vdw), at.q_base, at.h_b]
    et = params.elements.get(tname)
    if et is not None:
        return [et.r_vdw, math.sqrt(et.e_vdw), et.q_uff, 0.0]
    return [1.5, 0.0, 0.0, 0.0]


def bond_order_from_types(ta, tb):
    """params.rs:617-623."""
    if ta.endswith("_1") or tb.endswith("_1"):
        return 3.0
    if ta.endswith("_2") and tb.endswith("_2"):
        return 2.0
    return 1.0


def uff_bond_length(ti: AtomType, tj: AtomType, ei: ElementType, ej: ElementType, bo: float):
    """params.rs:627-634."""
    r_bo = -0.1332 * (ti.r_uff + tj.r_uff) * math.log(bo)
    if ei.e_aff < 0.0 and ej.e_aff < 0.0:
        s = math.sqrt(-ei.e_aff) - math.sqrt(-ej.e_aff)
        r_en = ti.r_uff * tj.r_uff * s * s / (-ei.e_aff * ti.r_uff - ej.e_aff * tj.r_uff)
    else:
        r_en = 0.0
    return ti.r_uff + tj.r_uff + r_bo - r_en


def uff_bond_k(ei: ElementType, ej: ElementType, l0: float):
    """params.rs:637-639."""
    return 0.5 * 28.7989689090648 * ei.q_uff * ej.q_uff / (l0 * l0 * l0)


def uff_angle_sp3(ct, st2):
    """params.rs:642-647."""
    c2 = 1.0 / (4.0 * st2)
    c1 = -4.0 * c2 * ct
    c0 = c2 * (2.0 * ct * ct + 1.0)
    return c0, c1, c2, 0.0


# ---------------------------------------------------------------------------
# topology  (moltopo/src/topology.rs, molecular.rs, assign_uff.rs)
# ---------------------------------------------------------------------------
def read_xyz(path):
    """moltopo/src/xyz.rs:12-35 (element, x, y, z[, q])."""
    lines = open(path).read().splitlines()
    n = int(lines[0].split()[0])
    elems, pos, q = [], [], []
    for ln in lines[2:2 + n]:
        w = ln.split()
        elems.append(w[0])
        pos.append([float(w[1]), float(w[2]), float(w[3])])
        q.append(float(w[4]) if len(w) > 4 else 0.0)
    return elems, np.array(pos, dtype=np.float64), np.array(q, dtype=np.float64)


def build_bonds_by_cutoff(apos, rcut):
    """topology.rs:101-111 (strict r^2 < rcut^2, i<j order)."""
    n = len(apos)
    r2cut = rcut * rcut
    bonds = []
    for i in range(n):
        for j in range(i + 1, n):
            d = apos[j] - apos[i]
            if float(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < r2cut:
                bonds.append((i, j))
    return bonds


def _adjacency(natoms, bonds):
    neigh = [[] for _ in range(natoms)]
    for a, b in bonds:
        neigh[a].append(b)
        neigh[b].append(a)
    return neigh


def build_angles_from_bonds(natoms, bonds):
    """topology.rs:113-129."""
    neigh = _adjacency(natoms, bonds)
    angles = []
    for j in range(natoms):
        ns = neigh[j]
        for a in range(len(ns)):
            for b in range(a + 1, len(ns)):
                angles.append((ns[a], j, ns[b]))
    return angles


def build_dihedrals_from_bonds(natoms, bonds):
    """topology.rs:131-159: the set {(i,j,k,l): j-k bonded, i~j, l~k, i!=k,
    l!=j, l>i}.  The reference returns it in HashSet order; this restatement
    returns it sorted lexicographically (order carries no meaning)."""
    neigh = _adjacency(natoms, bonds)
    s = set()
    for j in range(natoms):
        for k in neigh[j]:
            for i in neigh[j]:
                if i == k:
                    continue
                for l in neigh[k]:
                    if l == j:
                        continue
                    if l > i:
                        s.add((i, j, k, l))
    return sorted(s)


def build_inversions_from_bonds(natoms, bonds):
    """topology.rs:161-182."""
    neigh = _adjacency(natoms, bonds)
    invs = []
    for i in range(natoms):
        ns = neigh[i]
        if len(ns) == 3:
            invs.append((i, ns[0], ns[1], ns[2]))
            invs.append((i, ns[2], ns[0], ns[1]))
            invs.append((i, ns[1], ns[2], ns[0]))
    return invs


def make_neigh_bs(natoms, bonds):
    """Atoms::make_neigh_bs (molecular.rs:34-78): slot = insertion order."""
    neighs = -np.ones((natoms, 4), dtype=np.int32)
    neigh_bs = -np.ones((natoms, 4), dtype=np.int32)
    for ib, (ia, ja) in enumerate(bonds):
        for s in range(4):
            if neighs[ia, s] < 0:
                neighs[ia, s] = ja
                neigh_bs[ia, s] = ib
                break
        for s in range(4):
            if neighs[ja, s] < 0:
                neighs[ja, s] = ia
                neigh_bs[ja, s] = ib
                break
    return neighs, neigh_bs


def _ne_pairs(el):
    return {"H": 0, "C": 0, "N": 1, "O": 2, "F": 3, "Si": 0, "P": 1, "S": 2, "Cl": 3}.get(el, 0)


def hybridization(el, n_neigh):
    """topology.rs:28-34."""
    if el == "H":
        return 1
    npi = 4 - _ne_pairs(el) - n_neigh
    if npi < 0:
        return 1
    return 3 if npi == 0 else (2 if npi == 1 else 1)


_KNOWN_TYPES = ["H_", "C_3", "C_R", "C_2", "C_1", "N_3", "N_R", "N_2", "N_1", "O_3", "O_R", "O_2", "O_1",
                "F", "Cl", "Si3", "P", "S"]


def assign_uff_types(elems, neighs):
    """moltopo/src/assign_uff.rs:8-118."""
    n = len(elems)
    types = list(elems)
    set_atom = [False] * n
    for ia in range(n):
        if set_atom[ia]:
            continue
        name = elems[ia]
        nb = 0
        for s in range(4):
            if neighs[ia][s] < 0:
                break
            nb += 1
        h = hybridization(name, nb)
        if name == "H":
            types[ia] = "H_"
            set_atom[ia] = True
            continue
        if name == "N" and nb == 3:
            n_o = sum(1 for k in range(3) if neighs[ia][k] >= 0 and elems[neighs[ia][k]] == "O")
            if n_o == 2:
                types[ia] = "N_R"
                set_atom[ia] = True
                for k in range(3):
                    ja = neighs[ia][k]
                    if ja >= 0 and elems[ja] == "O":
                        types[ja] = "O_R"
                        set_atom[ja] = True
                continue
        if name == "O" and nb == 1:
            types[ia] = "O_2"
            set_atom[ia] = True
            continue
        if name == "C" and nb == 2:
            types[ia] = "C_1"
            set_atom[ia] = True
            continue
        if h == 3:
            suf = "3"
        elif h == 2:
            suf = "R" if name in ("C", "N", "O") else "2"
        elif h == 1:
            suf = "1"
        else:
            suf = "3"
        tname = f"{name}_{suf}"
        if tname in _KNOWN_TYPES:
            types[ia] = tname
            set_atom[ia] = True
        elif name in _KNOWN_TYPES:
            types[ia] = name
            set_atom[ia] = True
    return types


def _elem_sym(t):
    i = t.find("_")
    return t[:i] if i >= 0 else t


def _hyb_suffix(t):
    i = t.rfind("_")
    if i >= 0 and i + 1 < len(t):
        return t[i + 1]
    return "3"


def _calc_rij(params, ta, tb, bo):
    ti, tj = params.atom_types.get(ta), params.atom_types.get(tb)
    ei, ej = params.element_of_atom_type(ta), params.element_of_atom_type(tb)
    if ti is None or ei is None or tj is None or ej is None:
        return 1.5
    return uff_bond_length(ti, tj, ei, ej, bo)


def _nneigh(neighs, i):
    c = 0
    for s in range(4):
        if neighs[i][s] < 0:
            break
        c += 1
    return c


def setup_params(params: Params, types, neighs, bonds, angles, dihedrals, inversions):
    """Uff::setup_params (molff/src/uff.rs:699-887)."""
    deg2rad = math.pi / 180.0
    bon = []
    for (a, b) in bonds:
        ta, tb = types[a], types[b]
        bo = bond_order_from_types(ta, tb)
        l0 = _calc_rij(params, ta, tb, bo)
        ei, ej = params.element_of_atom_type(ta), params.element_of_atom_type(tb)
        if ei is None or ej is None:
            bon.append([100.0, l0])
            continue
        bon.append([uff_bond_k(ei, ej, l0), l0])
    ang = []
    for (i, j, k) in angles:
        tj, ti, tk = types[j], types[i], types[k]
        if _elem_sym(tj) == "H":
            ang.append([0.0] * 5)
            continue
        at_j = params.atom_types.get(tj)
        ei, ek = params.element_of_atom_type(ti), params.element_of_atom_type(tk)
        if at_j is None or ei is None or ek is None:
            ang.append([0.0] * 5)
            continue
        ct = math.cos(at_j.a_ss * deg2rad)
        st2 = math.sin(at_j.a_ss * deg2rad) ** 2
        rij = _calc_rij(params, ti, tj, bond_order_from_types(ti, tj))
        rjk = _calc_rij(params, tj, tk, bond_order_from_types(tj, tk))
        rik = math.sqrt(rij * rij + rjk * rjk - 2.0 * rij * rjk * ct)
        kappa = 28.7989689090648 * ei.q_uff * ek.q_uff / (rik ** 5) * (3.0 * rij * rjk * st2 - rik * rik * ct)
        hyb = _hyb_suffix(tj)
        if hyb == "1":
            row = [kappa, 1.0, 1.0, 0.0, 0.0]
        elif hyb in ("2", "R"):
            row = [kappa / 9.0, 1.0, 0.0, 0.0, -1.0]
        elif hyb == "3":
            c0, c1, c2, c3 = uff_angle_sp3(ct, st2)
            row = [kappa, c0, c1, c2, c3]
        else:
            row = [kappa, 1.0, 0.0, 0.0, -1.0]
        ang.append(row)
    dih = []
    for (i1, i2, i3, i4) in dihedrals:
        t2, t3 = types[i2], types[i3]
        if _elem_sym(t2) == "H" or _elem_sym(t3) == "H":
            dih.append([0.0, 1.0, 3.0])
            continue
        h2, h3 = _hyb_suffix(t2), _hyb_suffix(t3)
        if h2 == "1" or h3 == "1":
            dih.append([0.0, 1.0, 3.0])
            continue
        s2, s3 = h2 == "3", h3 == "3"
        e2, e3 = params.element_of_atom_type(t2), params.element_of_atom_type(t3)
        bo23 = bond_order_from_types(t2, t3)
        sp2 = ("2", "R")
        if s2 and s3:
            v = math.sqrt((e2.v_uff if e2 else 0.0) * (e3.v_uff if e3 else 0.0))
            el2, el3 = _elem_sym(t2), _elem_sym(t3)
            if el2 in ("O", "S") and el3 in ("O", "S"):
                k = KCAL_TO_EV
                k *= 2.0 if el2 == "O" else 6.8
                k *= 2.0 if el3 == "O" else 6.8
                v, d, n = math.sqrt(k), 1.0, 2.0
            else:
                d, n = 1.0, 3.0
        elif (s2 and h3 in sp2) or (h2 in sp2 and s3):
            v, d, n = KCAL_TO_EV, -1.0, 6.0
            el_sp3 = _elem_sym(t2) if s2 else _elem_sym(t3)
            if el_sp3 in ("O", "S"):
                v = 5.0 * math.sqrt((e2.u_uff if e2 else 0.0) * (e3.u_uff if e3 else 0.0)) * (1.0 + 4.18 * math.log(bo23))
                d, n = 1.0, 2.0
            sp3_idx, sp2_idx = (i2, i3) if s2 else (i3, i2)
            has = False
            for s in range(4):
                nn = neighs[sp2_idx][s]
                if nn < 0:
                    break
                if nn == sp3_idx:
                    continue
                if _hyb_suffix(types[nn]) in sp2:
                    has = True
                    break
            if has:
                v, d, n = 2.0 * KCAL_TO_EV, 1.0, 3.0
        elif h2 in sp2 and h3 in sp2:
            v = 5.0 * math.sqrt((e2.u_uff if e2 else 0.0) * (e3.u_uff if e3 else 0.0)) * (1.0 + 4.18 * math.log(bo23))
            d, n = -1.0, 2.0
        else:
            v, d, n = 0.0, 1.0, 3.0
        nb2, nb3 = float(_nneigh(neighs, i2)), float(_nneigh(neighs, i3))
        denom = 0.5 * (nb2 - 1.0) * (nb3 - 1.0)
        dih.append([v / denom if denom > 0.0 else v, d, n])
    inv = []
    for (i1, _j, _k, _l) in inversions:
        t1 = types[i1]
        el1, h1 = _elem_sym(t1), _hyb_suffix(t1)
        if el1 == "C" and h1 in ("2", "R"):
            carbonyl = False
            for s in range(4):
                nn = neighs[i1][s]
                if nn < 0:
                    break
                tn = types[nn]
                if _elem_sym(tn) == "O" and _hyb_suffix(tn) == "2":
                    carbonyl = True
                    break
            k = (50.0 if carbonyl else 6.0) * KCAL_TO_EV
            row = (k, 1.0, -1.0, 0.0)
        elif el1 == "N" and h1 in ("2", "R"):
            row = (6.0 * KCAL_TO_EV, 1.0, -1.0, 0.0)
        elif el1 == "N" and h1 == "3":
            row = (0.0, 0.0, 0.0, 0.0)
        elif el1 == "P":
            om = 84.4339 * deg2rad
            c0 = 4.0 * math.cos(om) ** 2 - math.cos(2.0 * om)
            c1 = -4.0 * math.cos(om)
            c2 = 1.0
            row = (22.0 * KCAL_TO_EV / (c0 + c1 + c2), c0, c1, c2)
        else:
            row = (0.0, 0.0, 0.0, 0.0)
        inv.append([row[0] / 3.0, row[1], row[2], row[3]])
    return (np.array(bon, dtype=np.float64).reshape(-1, 2), np.array(ang, dtype=np.float64).reshape(-1, 5),
            np.array(dih, dtype=np.float64).reshape(-1, 3), np.array(inv, dtype=np.float64).reshape(-1, 4))


def bake_angle_neighs(angles, neighs):
    """uff.rs:250-267."""
    out = []
    for (i, j, k) in angles:
        ji = jk = -1
        for s in range(4):
            ing = neighs[j][s]
            if ing < 0:
                break
            if ing == i:
                ji = j * 4 + s
            elif ing == k:
                jk = j * 4 + s
        out.append((ji, jk))
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def bake_dihedral_neighs(dihedrals, neighs):
    """uff.rs:269-294."""
    out = []
    for (i, j, k, l) in dihedrals:
        ji = jk = kl = -1
        for s in range(4):
            ing = neighs[j][s]
            if ing < 0:
                break
            if ing == i:
                ji = j * 4 + s
            elif ing == k:
                jk = j * 4 + s
        for s in range(4):
            ing = neighs[k][s]
            if ing < 0:
                break
            if ing == l:
                kl = k * 4 + s
        out.append((ji, jk, kl))
    return np.array(out, dtype=np.int32).reshape(-1, 3)


def bake_inversion_neighs(inversions, neighs):
    """uff.rs:296-315 (only slots 0..2 are scanned, as in the reference)."""
    out = []
    for (i, j, k, l) in inversions:
        ij = ik = il = -1
        for s in range(3):
            ing = neighs[i][s]
            if ing == j:
                ij = i * 4 + s
            elif ing == k:
                ik = i * 4 + s
            elif ing == l:
                il = i * 4 + s
        out.append((ij, ik, il))
    return np.array(out, dtype=np.int32).reshape(-1, 3)


def make_second_neighs(neighs):
    """NonBondedFF::make_second_neighs (nonbonded.rs:112-137) via the C oracle."""
    n = len(neighs)
    ng = _i32(neighs, (n, 4))
    ex = np.full((n, EXCL_MAX), -1, dtype=np.int32)
    rc = lib().orc_make_second_neighs(n, _pi(ng), _pi(ex))
    if rc != 0:
        raise ValueError(f"make_second_neighs: atom {-1 - rc} has more than {EXCL_MAX} exclusions")
    return ex


# ---------------------------------------------------------------------------
# a system bundle
# ---------------------------------------------------------------------------
@dataclass
class OracleSystem:
    """All per-system tables the FP64 oracle needs (SoA numpy arrays)."""

    elems: list
    apos: np.ndarray
    bonds: np.ndarray
    angles: np.ndarray
    dihedrals: np.ndarray
    inversions: np.ndarray
    neighs: np.ndarray
    neigh_bs: np.ndarray
    types: list
    bon_params: np.ndarray
    ang_params: np.ndarray
    dih_params: np.ndarray
    inv_params: np.ndarray
    ang_ngs: np.ndarray
    dih_ngs: np.ndarray
    inv_ngs: np.ndarray
    reqs: np.ndarray
    nb_mode: int = NB_GPU
    rdamp: float = 0.1
    fmax_nb: float = 10.0
    clamp_nb: int = 1
    rcut: float = 0.0
    excl: np.ndarray | None = None
    grid: np.ndarray | None = None
    grid_ns: tuple = (0, 0, 0)
    grid_p0: tuple = (0.0, 0.0, 0.0)
    grid_step: tuple = (1.0, 1.0, 1.0)
    alpha_morse: float = 1.5
    mol_id: np.ndarray | None = None      # coupled layer: molecule of each atom
    rcut_inter: float = 0.0               # coupled layer: cutoff of pairs between molecules
    layer_cells: bool = False             # coupled layer: cell-grid pair search (same terms, O(N))
    rswitch_inter: float = 0.0            # coupled layer: width of the smooth switch below the cutoff
    pbc: tuple = (0.0, 0.0)               # coupled layer: lateral periodic box (Lx, Ly); 0 = open
    _keep: list = field(default_factory=list)

    @property
    def natoms(self):
        return len(self.elems)

    def cstruct(self) -> OrcSystem:
        s = OrcSystem()
        arrs = {
            "neighs": _i32(self.neighs, (-1,)), "neigh_bs": _i32(self.neigh_bs, (-1,)),
            "bon_params": _f64(self.bon_params, (-1,)),
            "ang_atoms": _i32(self.angles, (-1,)), "ang_ngs": _i32(self.ang_ngs, (-1,)),
            "ang_params": _f64(self.ang_params, (-1,)),
            "dih_atoms": _i32(self.dihedrals, (-1,)), "dih_ngs": _i32(self.dih_ngs, (-1,)),
            "dih_params": _f64(self.dih_params, (-1,)),
            "inv_atoms": _i32(self.inversions, (-1,)), "inv_ngs": _i32(self.inv_ngs, (-1,)),
            "inv_params": _f64(self.inv_params, (-1,)),
            "reqs": _f64(self.reqs, (-1,)),
        }
        # zero-length arrays still need a valid pointer
        for k, v in list(arrs.items()):
            if v.size == 0:
                arrs[k] = np.zeros(1, dtype=v.dtype)
        self._keep = list(arrs.values())
        s.natoms = self.natoms
        s.nbonds = len(self.bonds)
        s.nangles = len(self.angles)
        s.ndih = len(self.dihedrals)
        s.ninv = len(self.inversions)
        for k, v in arrs.items():
            setattr(s, k, _pi(v) if v.dtype == np.int32 else _pd(v))
        s.nb_mode = int(self.nb_mode)
        if self.nb_mode == NB_CPU:
            ex = self.excl if self.excl is not None else make_second_neighs(self.neighs)
            ex = _i32(ex, (-1,))
            self._keep.append(ex)
            s.excl = _pi(ex)
        s.rdamp = self.rdamp
        s.fmax_nb = self.fmax_nb
        s.clamp_nb = int(self.clamp_nb)
        s.rcut2 = self.rcut * self.rcut if self.rcut > 0 else 1e300
        if self.grid is not None:
            g = _f64(self.grid, (-1,))
            self._keep.append(g)
            s.grid = _pd(g)
            s.gnx, s.gny, s.gnz = (int(x) for x in self.grid_ns)
            for k in range(3):
                s.gp0[k] = self.grid_p0[k]
                s.ginv_step[k] = 1.0 / self.grid_step[k]
            s.alpha_morse = self.alpha_morse
        if self.mol_id is not None:
            m = _i32(self.mol_id, (-1,))
            self._keep.append(m)
            s.mol_id = _pi(m)
            s.rcut_inter2 = self.rcut_inter * self.rcut_inter
            s.layer_cells = 1 if self.layer_cells else 0
            s.rswitch_inter = self.rswitch_inter
            s.pbc_lx, s.pbc_ly = float(self.pbc[0]), float(self.pbc[1])
        return s


def system_from_geometry(elems, apos, params: Params, charges=None, rcut_bond=1.8, bonds=None, **kw) -> OracleSystem:
    """The reference's set-up sequence (relax_pentacene_uff.rs:53-78 and
    molengine setup_nonbonded main.rs:313-337), restated."""
    apos = np.asarray(apos, dtype=np.float64)
    n = len(elems)
    if bonds is None:
        bonds = build_bonds_by_cutoff(apos, rcut_bond)
    bonds = [tuple(int(x) for x in b) for b in bonds]
    angles = build_angles_from_bonds(n, bonds)
    dihedrals = build_dihedrals_from_bonds(n, bonds)
    inversions = build_inversions_from_bonds(n, bonds)
    neighs, neigh_bs = make_neigh_bs(n, bonds)
    types = assign_uff_types(elems, neighs.tolist())
    bp, ap, dp, ip = setup_params(params, types, neighs.tolist(), bonds, angles, dihedrals, inversions)
    reqs = np.array([get_reqh(params, t) for t in types], dtype=np.float64)
    if charges is not None:
        reqs[:, 2] = np.asarray(charges, dtype=np.float64)
    nl = neighs.tolist()
    return OracleSystem(
        elems=list(elems), apos=apos.copy(), bonds=np.array(bonds, dtype=np.int32).reshape(-1, 2),
        angles=np.array(angles, dtype=np.int32).reshape(-1, 3),
        dihedrals=np.array(dihedrals, dtype=np.int32).reshape(-1, 4),
        inversions=np.array(inversions, dtype=np.int32).reshape(-1, 4),
        neighs=neighs, neigh_bs=neigh_bs, types=types,
        bon_params=bp, ang_params=ap, dih_params=dp, inv_params=ip,
        ang_ngs=bake_angle_neighs(angles, nl), dih_ngs=bake_dihedral_neighs(dihedrals, nl),
        inv_ngs=bake_inversion_neighs(inversions, nl), reqs=reqs, **kw)


def system_from_terms(types, apos, bonds, angles, dihedrals, inversions, params: Params, charges=None,
                      **kw) -> OracleSystem:
    """A system from given term lists, kept in their order: Uff::from_topology
    (molff/src/uff.rs:238-240) takes the topology's bonds, angles, dihedrals
    and inversions as they are, then the set-up of relax_pentacene_uff.rs:
    53-78 follows (neighbour tables, setup_params with the given UFF types,
    hneigh slots, REQH).  Used for topologies read from the reference's file
    formats, whose dihedrals come in the writer's HashSet order."""
    apos = np.asarray(apos, dtype=np.float64)
    n = len(types)
    bonds = [tuple(int(x) for x in b) for b in bonds]
    angles = [tuple(int(x) for x in a) for a in angles]
    dihedrals = [tuple(int(x) for x in d) for d in dihedrals]
    inversions = [tuple(int(x) for x in v) for v in inversions]
    neighs, neigh_bs = make_neigh_bs(n, bonds)
    nl = neighs.tolist()
    bp, ap, dp, ip = setup_params(params, list(types), nl, bonds, angles, dihedrals, inversions)
    reqs = np.array([get_reqh(params, t) for t in types], dtype=np.float64)
    if charges is not None:
        reqs[:, 2] = np.asarray(charges, dtype=np.float64)
    elems = [params.atom_types[t].element if t in params.atom_types else _elem_sym(t) for t in types]
    return OracleSystem(
        elems=elems, apos=apos.copy(), bonds=np.array(bonds, dtype=np.int32).reshape(-1, 2),
        angles=np.array(angles, dtype=np.int32).reshape(-1, 3),
        dihedrals=np.array(dihedrals, dtype=np.int32).reshape(-1, 4),
        inversions=np.array(inversions, dtype=np.int32).reshape(-1, 4),
        neighs=neighs, neigh_bs=neigh_bs, types=list(types),
        bon_params=bp, ang_params=ap, dih_params=dp, inv_params=ip,
        ang_ngs=bake_angle_neighs(angles, nl), dih_ngs=bake_dihedral_neighs(dihedrals, nl),
        inv_ngs=bake_inversion_neighs(inversions, nl), reqs=reqs, **kw)


def layer_from_molecules(mols, positions, params: Params, rcut_inter: float, **kw) -> OracleSystem:
    """A coupled layer: the union of several molecules as one system whose
    pairs between different molecules interact within rcut_inter only (the
    GPU pair convention inside each molecule).  mols: objects with elems,
    bonds, charges; positions: one [n,3] array per molecule."""
    elems, apos, bonds, charges, mol_id = [], [], [], [], []
    off = 0
    for k, (m, x) in enumerate(zip(mols, positions)):
        elems += list(m.elems)
        apos.append(np.asarray(x, dtype=np.float64))
        bonds += [(int(a) + off, int(b) + off) for a, b in np.asarray(m.bonds).reshape(-1, 2)]
        charges.append(np.asarray(m.charges, dtype=np.float64))
        mol_id += [k] * len(m.elems)
        off += len(m.elems)
    return system_from_geometry(elems, np.concatenate(apos), params, charges=np.concatenate(charges), bonds=bonds,
                                nb_mode=NB_GPU, mol_id=np.array(mol_id, dtype=np.int32), rcut_inter=rcut_inter, **kw)


# ---------------------------------------------------------------------------
# thin wrappers around the C oracle
# ---------------------------------------------------------------------------
def uff_eval(sys: OracleSystem, apos=None):
    cs = sys.cstruct()
    x = _f64(sys.apos if apos is None else apos, (-1, 3))
    f = np.zeros_like(x)
    E = np.zeros(4)
    lib().orc_uff_eval(ctypes.byref(cs), _pd(x), _pd(f), _pd(E))
    return f, E


def nb_eval(sys: OracleSystem, apos=None):
    cs = sys.cstruct()
    x = _f64(sys.apos if apos is None else apos, (-1, 3))
    f = np.zeros_like(x)
    e = lib().orc_nb_eval(ctypes.byref(cs), _pd(x), _pd(f))
    return f, e


def grid_eval(sys: OracleSystem, apos=None):
    cs = sys.cstruct()
    x = _f64(sys.apos if apos is None else apos, (-1, 3))
    f = np.zeros_like(x)
    e = lib().orc_grid_eval(ctypes.byref(cs), _pd(x), _pd(f))
    return f, e


def force_scale(sys: OracleSystem, apos=None):
    """Per atom, the sum of the magnitudes of every force contribution the
    atom's total is summed from (bond, each angle/dihedral/inversion piece,
    each pair, the surface): the scale against which the rounding error of
    that sum is measured.  Near a minimum the total goes to zero while these
    contributions stay of order eV/A."""
    cs = sys.cstruct()
    x = _f64(sys.apos if apos is None else apos, (-1, 3))
    scale = np.zeros(len(x))
    cs.fscale = _pd(scale)
    f = np.zeros_like(x)
    E = np.zeros(6)
    lib().orc_eval_forces(ctypes.byref(cs), _pd(x), _pd(f), _pd(E))
    return scale


def eval_forces(sys: OracleSystem, apos=None):
    """Returns (F[N,3], E6) with E6 = (bond, angle, dihedral, inversion, nb, surface)."""
    cs = sys.cstruct()
    x = _f64(sys.apos if apos is None else apos, (-1, 3))
    f = np.zeros_like(x)
    E = np.zeros(6)
    lib().orc_eval_forces(ctypes.byref(cs), _pd(x), _pd(f), _pd(E))
    return f, E


@dataclass
class FireParams:
    dt0: float
    dt_max: float
    alpha0: float = 0.1
    n_min: int = 5
    f_inc: float = 1.1
    f_dec: float = 0.5
    f_alpha: float = 0.99
    flim: float = 0.0

    def cstruct(self):
        return OrcFire(self.dt0, self.dt_max, self.alpha0, self.n_min, self.f_inc, self.f_dec, self.f_alpha,
                       self.flim)


def relax_fire(sys: OracleSystem, fire: FireParams, max_steps: int, fconv: float, apos=None, vapos=None):
    """FP64 FIRE relaxation.  Returns dict(steps, E, fmax, E6, converged, apos, vapos)."""
    cs = sys.cstruct()
    x = _f64(sys.apos if apos is None else apos, (-1, 3)).copy()
    v = np.zeros_like(x) if vapos is None else _f64(vapos, (-1, 3)).copy()
    out = np.zeros(9)
    fc = fire.cstruct()
    steps = lib().orc_relax_fire(ctypes.byref(cs), _pd(x), _pd(v), ctypes.byref(fc), int(max_steps), float(fconv),
                                 _pd(out))
    return {"steps": steps, "E": out[0], "fmax": out[1], "E6": out[2:8].copy(), "converged": bool(out[8]),
            "apos": x, "vapos": v}


def run_md(sys: OracleSystem, niter, dt, fconv, flim, damping, apos=None, vapos=None):
    cs = sys.cstruct()
    x = _f64(sys.apos if apos is None else apos, (-1, 3)).copy()
    v = np.zeros_like(x) if vapos is None else _f64(vapos, (-1, 3)).copy()
    steps = lib().orc_run_md(ctypes.byref(cs), _pd(x), _pd(v), int(niter), float(dt), float(fconv), float(flim),
                             float(damping))
    return steps, x, v


def fire_step(forces, apos, vapos, state, fire: FireParams):
    """One FIRE step (orc_fire_step) on given forces [N,3].  state = [dt,
    alpha, n_pos] is updated in place.  Returns (apos, vapos, clamped forces,
    max_i |F_i|^2 of the given forces)."""
    f = _f64(forces, (-1, 3)).copy()
    x = _f64(apos, (-1, 3)).copy()
    v = _f64(vapos, (-1, 3)).copy()
    st = np.ascontiguousarray(state, dtype=np.float64)
    fc = fire.cstruct()
    f2max = lib().orc_fire_step(len(x), _pd(f), _pd(v), _pd(x), _pd(st), ctypes.byref(fc))
    state[:] = st
    return x, v, f, f2max


def md_step(forces, apos, vapos, dt, flim, cdamp):
    """One damped-MD step of run_md (orc_md_step) on given forces [N,3].
    Returns (apos, vapos, (sum v.F, sum |v|^2, sum |F|^2))."""
    f = _f64(forces, (-1, 3)).copy()
    x = _f64(apos, (-1, 3)).copy()
    v = _f64(vapos, (-1, 3)).copy()
    sums = np.zeros(3)
    lib().orc_md_step(len(x), _pd(f), _pd(v), _pd(x), float(dt), float(flim), float(cdamp), _pd(sums))
    return x, v, tuple(sums)


def md_steps(sys: OracleSystem, nsteps, dt, flim, cdamp, apos, vapos):
    """Plain move_atom_md steps (no purge): returns (apos, vapos)."""
    cs = sys.cstruct()
    x = _f64(apos, (-1, 3)).copy()
    v = _f64(vapos, (-1, 3)).copy()
    lib().orc_md_steps(ctypes.byref(cs), _pd(x), _pd(v), int(nsteps), float(dt), float(flim), float(cdamp))
    return x, v


def fe3d_pbc_comb(u, grid, ns, plqh):
    g = _f64(grid, (-1,))
    uu = _f64(u, (3,))
    p = _f64(plqh, (4,))
    out = np.zeros(4)
    lib().orc_fe3d_pbc_comb(_pd(uu), int(ns[0]), int(ns[1]), int(ns[2]), _pd(g), _pd(p), _pd(out))
    return out


def plqh(req, alpha):
    r = _f64(req, (4,))
    out = np.zeros(4)
    lib().orc_plqh(_pd(r), float(alpha), _pd(out))
    return out


def gridff_build(sub_pos, sub_reqs, lvec, ns, p0, npbc, alpha, rdamp):
    """Reference grid (Pauli, London, Coulomb, 0) at every node, z-fastest."""
    sp = _f64(sub_pos, (-1, 3))
    sr = _f64(sub_reqs, (-1, 4))
    lv = _f64(lvec, (9,))
    p = _f64(p0, (3,))
    nb = _i32(npbc, (3,))
    nx, ny, nz = (int(x) for x in ns)
    out = np.zeros(nx * ny * nz * 4)
    lib().orc_gridff_build(len(sp), _pd(sp), _pd(sr), _pd(lv), nx, ny, nz, _pd(p), _pi(nb), float(alpha),
                           float(rdamp), _pd(out))
    return out.reshape(nx, ny, nz, 4)


def gridff_build_nodes(sub_pos, sub_reqs, lvec, ns, p0, npbc, alpha, rdamp, nodes):
    """Reference-grid values at the listed nodes [(ix, iy, iz), ...] -> [n, 4]."""
    sp = _f64(sub_pos, (-1, 3))
    sr = _f64(sub_reqs, (-1, 4))
    nd = _i32(nodes, (-1, 3))
    out = np.zeros((len(nd), 4))
    nx, ny, nz = (int(x) for x in ns)
    lib().orc_gridff_build_nodes(len(sp), _pd(sp), _pd(sr), _pd(_f64(lvec, (9,))), nx, ny, nz, _pd(_f64(p0, (3,))),
                                 _pi(_i32(npbc, (3,))), float(alpha), float(rdamp), len(nd), _pi(nd), _pd(out))
    return out


def bspline_apply(C, ns):
    nx, ny, nz = (int(x) for x in ns)
    c = _f64(C, (-1,))
    out = np.zeros_like(c)
    lib().orc_bspline_apply(nx, ny, nz, _pd(c), _pd(out))
    return out.reshape(nx, ny, nz, 4)


def bspline_fit(E, ns):
    nx, ny, nz = (int(x) for x in ns)
    e = _f64(E, (-1,))
    out = np.zeros_like(e)
    lib().orc_bspline_fit(nx, ny, nz, _pd(e), _pd(out))
    return out.reshape(nx, ny, nz, 4)


# ---------------------------------------------------------------------------
# multi-replica rigid-atom port relaxation (RAFF "RRsp3" cluster step)
# ---------------------------------------------------------------------------
_RAFF_MASS = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999, "F": 18.998, "P": 30.974, "S": 32.06, "Cl": 35.45,
              "Br": 79.904, "I": 126.904, "Na": 22.990, "K": 39.098, "Mg": 24.305, "Ca": 40.078, "Fe": 55.845,
              "Cu": 63.546, "Zn": 65.38, "Al": 26.982, "Si": 28.085, "B": 10.81, "Li": 6.94, "Be": 9.012,
              "Ne": 20.180, "Ar": 39.948}


def raff_pack(elems, pos, bonds, nnode=-1, group_size=0, k_stiff=200.0):
    """One replica's tables, restating oclff::pack for one molecule
    (crates/libs/oclff/src/pack.rs): pack_molecules (:54-116, nodes = bond
    degree > 1 first, caps, padding "X" to group_size), masses_from_elems
    (:132-150, float32 inverse masses), build_neighs_from_bonds (:155-170),
    make_exclusions_1st_2nd (:176-211), make_ports_from_neighs (:293-316) and
    make_bk_slots_clustered (:218-244).  Float tables are float32 as there."""
    n = len(elems)
    pos = np.asarray(pos, dtype=np.float32).reshape(n, 3)
    N = group_size or n
    if n > N:
        raise ValueError("more atoms than group_size")
    deg = [0] * n
    for i, j in bonds:
        if i != j:
            deg[i] += 1
            deg[j] += 1
    ninf = sum(1 for x in deg if x > 1)
    if nnode < 0:
        nnode = ninf
    if nnode != ninf:
        raise ValueError(f"nnode {nnode} but {ninf} atoms of degree > 1")
    perm = [i for i in range(n) if deg[i] > 1] + [i for i in range(n) if deg[i] <= 1]
    inv = {p: k for k, p in enumerate(perm)}
    el = [elems[p] for p in perm] + ["X"] * (N - n)
    P = np.zeros((N, 3), dtype=np.float32)
    P[:n] = pos[perm]
    inv_mass = np.zeros(N, dtype=np.float32)
    for k in range(n):
        inv_mass[k] = np.float32(1.0) / np.float32(_RAFF_MASS.get(el[k], 12.0))
    pb = [(inv[i], inv[j]) for i, j in bonds]
    neighs = -np.ones((N, 4), dtype=np.int32)
    d = [0] * N
    for i, j in pb:
        if d[i] >= 4 or d[j] >= 4:
            raise ValueError("more than 4 bonds")
        neighs[i, d[i]] = j
        neighs[j, d[j]] = i
        d[i] += 1
        d[j] += 1
    excl1 = -np.ones((N, 4), dtype=np.int32)
    excl2 = -np.ones((N, 4), dtype=np.int32)
    for i in range(N):
        n1 = [int(x) for x in neighs[i] if x >= 0]
        excl1[i, :len(n1)] = n1
        s2 = []
        for j in n1:
            for t in neighs[j]:
                t = int(t)
                if t < 0 or t == i or t in n1 or t in s2:
                    continue
                s2.append(t)
                if len(s2) >= 4:
                    break
            if len(s2) >= 4:
                break
        excl2[i, :len(s2)] = s2
    port = np.zeros((N, 4, 3), dtype=np.float32)
    kflat = np.zeros((N, 4), dtype=np.float32)
    for i in range(N):
        for k in range(4):
            j = neighs[i, k]
            if j >= 0:
                port[i, k] = P[j] - P[i]
                kflat[i, k] = k_stiff
    bk = -np.ones((N, 4), dtype=np.int32)
    cnt = [0] * N
    for il in range(nnode):
        for k in range(4):
            ja = neighs[il, k]
            if ja < 0:
                continue
            if cnt[ja] >= 4:
                raise ValueError("more than 4 back slots")
            bk[ja, cnt[ja]] = 4 * il + k
            cnt[ja] += 1
    return dict(natoms=N, nreal=n, nnode=nnode, perm=np.array(perm, dtype=np.int32), elems=el, pos3=P,
                inv_mass=inv_mass, neighs=neighs, excl1=excl1, excl2=excl2, port_local=port, kflat=kflat, bk_slots=bk)


class RaffReplica:
    """One replica's shared topology for orc_raff_step (FP64 copies of the
    float32 tables)."""

    def __init__(self, tab, radius=None, fixmask=None):
        N, nn = tab["natoms"], tab["nnode"]
        self.n, self.nnode = N, nn
        self.radius = np.zeros(N) if radius is None else np.asarray(radius, dtype=np.float64).copy()
        self.neighs = _i32(tab["neighs"], (-1,))
        self.excl1 = _i32(tab["excl1"], (-1,))
        self.excl2 = _i32(tab["excl2"], (-1,))
        self.bk_slots = _i32(tab["bk_slots"], (-1,))
        self.fixmask = None if fixmask is None else _i32(fixmask, (-1,))
        self.port_local = _f64(np.asarray(tab["port_local"], dtype=np.float64)[:nn].reshape(-1), (-1,))
        self.kflat = _f64(np.asarray(tab["kflat"], dtype=np.float64)[:nn].reshape(-1), (-1,))
        if self.port_local.size == 0:
            self.port_local = np.zeros(1)
            self.kflat = np.zeros(1)
        self.inv_mass = _f64(tab["inv_mass"], (-1,))
        self.collide = bool((self.radius > 0).any())
        s = OrcRaff()
        s.n, s.nnode = N, nn
        s.radius, s.neighs, s.excl1, s.excl2 = _pd(self.radius), _pi(self.neighs), _pi(self.excl1), _pi(self.excl2)
        s.bk_slots = _pi(self.bk_slots)
        s.fixmask = _pi(self.fixmask) if self.fixmask is not None else None
        s.port_local, s.kflat, s.inv_mass = _pd(self.port_local), _pd(self.kflat), _pd(self.inv_mass)
        self.cs = s

    def run(self, pos, quat=None, nsteps=1, dt=0.1, relaxation=0.5, beta=0.0, rot_mass_scale=1.0, tol=None,
            mom=None, qmom=None):
        """nsteps steps (or until the step's largest correction component is
        below tol).  Returns dict(pos, quat, mom, qmom, dcoll, dnode, drot,
        dneigh, metric, steps)."""
        N, nn = self.n, self.nnode
        x = _f64(pos, (N, 3)).copy()
        q = np.tile([0.0, 0.0, 0.0, 1.0], (N, 1)) if quat is None else _f64(quat, (N, 4)).copy()
        m = np.zeros((N, 3)) if mom is None else _f64(mom, (N, 3)).copy()
        qm = np.zeros((N, 4)) if qmom is None else _f64(qmom, (N, 4)).copy()
        dc = np.zeros((N, 3))
        dn = np.zeros((max(nn, 1), 3))
        dr = np.zeros((max(nn, 1), 3))
        dg = np.zeros((max(nn, 1) * 4, 3))
        metric, steps = 0.0, 0
        for _ in range(nsteps):
            metric = lib().orc_raff_step(ctypes.byref(self.cs), float(dt), float(rot_mass_scale), float(relaxation),
                                         float(beta), 1 if self.collide else 0, _pd(x), _pd(q), _pd(m), _pd(qm),
                                         _pd(dc), _pd(dn), _pd(dr), _pd(dg))
            steps += 1
            if tol is not None and metric < tol:
                break
        return dict(pos=x, quat=q, mom=m, qmom=qm, dcoll=dc, dnode=dn[:nn], drot=dr[:nn], dneigh=dg[:4 * nn],
                    metric=metric, steps=steps)
