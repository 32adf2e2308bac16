"""Seeded synthetic workloads for the relaxation path.

Everything here is deterministic given a seed (numpy PCG64 streams), so the
same inputs can be handed to the HIP library, to the FP64 checker and to the
CPU baseline.

* ``ptcda_like``       -- the 38-atom C24H8O6 planar molecule: perylene core +
                          two anhydride rings laid on a honeycomb lattice with
                          1.40 A ring bonds, 1.09 A C-H, 1.22 A C=O.
* ``random_polyhex``   -- random planar fused 6-ring graphs with C/N/O/H.
* ``rocksalt_slab``    -- 2-layer rock-salt (001) slab, +-0.9 e.
* ``place_on_surface`` -- in-plane rotation + xy offset + height.
* ``layer_lattice``    -- a coupled monolayer of PTCDA-like molecules.
* ``poisson_arrivals`` -- Poisson request stream for the serving loop.

Geometric conventions (chosen here, stated once):

* The rock-salt lattice constant is strained from 5.64 A to 6.00 A so that
  every periodic cell used (24, 48, 180, 360 A) holds an integer number of
  lattice periods, and every period an integer number (60) of 0.1 A grid
  steps.  The top layer lies in the plane z = 0, the second at z = -a/2.
* Molecule heights are measured from the top layer to the molecule's
  centroid.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SEED_BASE = 0x5EED0000
NACL_A = 6.0          # strained rock-salt lattice constant [A]
NACL_Q = 0.9          # ionic charge [e]
B_RING = 1.40         # ring bond [A]
B_CH = 1.09
B_CO = 1.22
FS_PER_TIME_UNIT = 10.1805054836  # sqrt(amu*A^2/eV) in fs; masses are 1 amu


def fs_to_internal(t_fs: float) -> float:
    """Convert femtoseconds to the internal time unit (eV, A, amu)."""
    return t_fs / FS_PER_TIME_UNIT


@dataclass
class Molecule:
    elems: list           # element symbols
    apos: np.ndarray      # [N,3] float64
    bonds: np.ndarray     # [nb,2] int32 (ring bonds then substituent bonds)
    charges: np.ndarray   # [N] float64

    @property
    def natoms(self):
        return len(self.elems)


# ---------------------------------------------------------------------------
# honeycomb helpers: hexagon centres on a triangular lattice (axial q,r)
# ---------------------------------------------------------------------------
_SQ3 = math.sqrt(3.0)


def _hex_center(q, r, b=B_RING):
    # centres spaced sqrt(3)*b; hexagon vertices at angles 0,60,..,300 deg
    return np.array([1.5 * b * q, _SQ3 * b * (r + 0.5 * q)])


def _hex_vertices(q, r, b=B_RING):
    c = _hex_center(q, r, b)
    ang = np.deg2rad(np.arange(6) * 60.0)
    return c[None, :] + b * np.stack([np.cos(ang), np.sin(ang)], axis=1)


_HEX_NEIGH = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, -1), (-1, 1)]


def _polyhex_graph(hexes, b=B_RING):
    """Vertices (deduplicated) and edges of the union of hexagons."""
    key2idx = {}
    verts = []
    edges = set()
    for (q, r) in hexes:
        vs = _hex_vertices(q, r, b)
        ids = []
        for v in vs:
            k = (int(round(v[0] * 1000)), int(round(v[1] * 1000)))
            if k not in key2idx:
                key2idx[k] = len(verts)
                verts.append(v)
            ids.append(key2idx[k])
        for a in range(6):
            i, j = ids[a], ids[(a + 1) % 6]
            edges.add((min(i, j), max(i, j)))
    verts = np.array(verts)
    # lattice neighbours that no hexagon of the set joins (they face each
    # other across an enclosed gap) are bonded too: the induced subgraph
    from scipy.spatial import cKDTree
    for i, j in cKDTree(verts).query_pairs(1.05 * b):
        edges.add((min(i, j), max(i, j)))
    return verts, sorted(edges)


def _outward_dir(vi, verts, nbrs):
    """Unit vector along the missing honeycomb bond of a 2-coordinated vertex."""
    d = np.zeros(2)
    for j in nbrs:
        u = verts[j] - verts[vi]
        d -= u / np.linalg.norm(u)
    return d / np.linalg.norm(d)


def _assemble(verts, edges, ring_elems, substituents):
    """substituents: {vertex: element} placed along the outward direction."""
    nv = len(verts)
    adj = [[] for _ in range(nv)]
    for i, j in edges:
        adj[i].append(j)
        adj[j].append(i)
    elems = list(ring_elems)
    pos = [np.array([v[0], v[1], 0.0]) for v in verts]
    bonds = [list(e) for e in edges]
    for vi in sorted(substituents):
        el = substituents[vi]
        d = _outward_dir(vi, verts, adj[vi])
        L = B_CH if el == "H" else B_CO
        pos.append(np.array([verts[vi][0] + L * d[0], verts[vi][1] + L * d[1], 0.0]))
        bonds.append([vi, len(elems)])
        elems.append(el)
    return elems, np.array(pos, dtype=np.float64), np.array(bonds, dtype=np.int32)


def _jitter_and_charge(rng, elems, pos, bonds, sigma_pos=0.05, sigma_q=0.2):
    pos = pos + rng.normal(0.0, sigma_pos, size=pos.shape)
    q = rng.normal(0.0, sigma_q, size=len(elems))
    q -= q.mean()
    return Molecule(list(elems), pos, bonds, q)


def ptcda_like(seed: int = SEED_BASE + 0, jitter: bool = True) -> Molecule:
    """C24H8O6: 7 fused rings (perylene core + 2 anhydride rings).

    Ring vertices lie on a honeycomb with 1.40 A bonds; the two anhydride
    bridge vertices are O; the four anhydride carbons carry =O (1.22 A); the
    remaining eight 2-coordinated carbons carry H (1.09 A).  Centred at the
    origin, molecular plane z = 0."""
    hexes = [(0, 0), (-1, 0), (-1, 1), (1, 0), (1, -1), (-2, 1), (2, -1)]
    verts, edges = _polyhex_graph(hexes)
    nv = len(verts)
    deg = np.zeros(nv, dtype=int)
    for i, j in edges:
        deg[i] += 1
        deg[j] += 1
    xmax = np.abs(verts[:, 0]).max()
    ring_elems = ["C"] * nv
    subs = {}
    for vi in range(nv):
        if deg[vi] != 2:
            continue
        x, y = verts[vi]
        if abs(abs(x) - xmax) < 1e-6 and abs(y) < 1e-6:
            ring_elems[vi] = "O"          # anhydride bridge
        elif abs(x) > xmax - 0.8:
            subs[vi] = "O"                # carbonyl
        else:
            subs[vi] = "H"
    elems, pos, bonds = _assemble(verts, edges, ring_elems, subs)
    pos[:, :2] -= pos[:, :2].mean(axis=0)
    rng = np.random.Generator(np.random.PCG64(seed))
    if not jitter:
        return Molecule(elems, pos, bonds, np.zeros(len(elems)))
    return _jitter_and_charge(rng, elems, pos, bonds)


def random_polyhex(rng: np.random.Generator, target_atoms: int) -> Molecule:
    """Random planar fused 6-ring molecule with about ``target_atoms`` atoms.

    Hexagons are added one at a time at random free neighbour positions of the
    current animal; the growth stops at the size whose atom count is closest
    to the target.  Vertex elements: 3-ring-coordinated vertices are C (N with
    probability 0.05); 2-coordinated vertices are C-H (0.80), C=O (0.07),
    pyridinic N (0.08) or ring O (0.05)."""
    from sortedcontainers import SortedList

    # incremental graph of the animal: vertex degrees, edge set, and the
    # free neighbour positions (kept sorted: the random pick indexes them)
    deg = {}
    edges = set()
    state = {"nv": 0, "ndeg2": 0}
    hexes, hexset, frontier = [], set(), SortedList()

    def add_hex(h):
        hexes.append(h)
        hexset.add(h)
        if h in frontier:
            frontier.remove(h)
        for dq, dr in _HEX_NEIGH:
            nb = (h[0] + dq, h[1] + dr)
            if nb not in hexset and nb not in frontier:
                frontier.add(nb)
        vs = _hex_vertices(h[0], h[1])
        keys = [(int(round(v[0] * 1000)), int(round(v[1] * 1000))) for v in vs]
        for k in keys:
            if k not in deg:
                deg[k] = 0
                state["nv"] += 1
        for a in range(6):
            e = tuple(sorted((keys[a], keys[(a + 1) % 6])))
            if e in edges:
                continue
            edges.add(e)
            for k in e:
                if deg[k] == 2:
                    state["ndeg2"] -= 1
                deg[k] += 1
                if deg[k] == 2:
                    state["ndeg2"] += 1

    def count():
        return state["nv"] + state["ndeg2"]   # upper estimate incl. substituents

    add_hex((0, 0))
    prev = None
    while True:
        c = count()
        if c >= target_atoms:
            if prev is not None and abs(prev[1] - target_atoms) < abs(c - target_atoms) and prev[1] >= 20:
                hexes = hexes[:prev[0]]
            break
        prev = (len(hexes), c)
        h = frontier[int(rng.integers(len(frontier)))]
        add_hex(h)
    verts, edges = _polyhex_graph(hexes)
    nv = len(verts)
    deg = np.zeros(nv, dtype=int)
    for i, j in edges:
        deg[i] += 1
        deg[j] += 1
    ring_elems = []
    subs = {}
    for vi in range(nv):
        u = rng.random()
        if deg[vi] == 3:
            ring_elems.append("N" if u < 0.05 else "C")
        else:
            if u < 0.80:
                ring_elems.append("C")
                subs[vi] = "H"
            elif u < 0.87:
                ring_elems.append("C")
                subs[vi] = "O"
            elif u < 0.95:
                ring_elems.append("N")
            else:
                ring_elems.append("O")
    # coves: two 2-coordinated vertices whose missing honeycomb neighbour is
    # the same lattice point would put their substituents 0.5 A apart; such
    # vertices carry no substituent and become pyridinic N
    adj = [[] for _ in range(nv)]
    for i, j in edges:
        adj[i].append(j)
        adj[j].append(i)
    missing = {}
    for vi in range(nv):
        if deg[vi] == 2:
            m = verts[vi] + B_RING * _outward_dir(vi, verts, adj[vi])
            missing.setdefault((int(round(m[0] * 100)), int(round(m[1] * 100))), []).append(vi)
    for group in missing.values():
        if len(group) > 1:
            for vi in group:
                subs.pop(vi, None)
                ring_elems[vi] = "N"
    elems, pos, bonds = _assemble(verts, edges, ring_elems, subs)
    pos[:, :2] -= pos[:, :2].mean(axis=0)
    return _jitter_and_charge(rng, elems, pos, bonds)


def uneven_batch(nsys: int = 2048, nmin: int = 20, nmax: int = 400, seed: int = SEED_BASE + 2):
    """Molecules with atom counts drawn log-uniform in [nmin, nmax]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mols = []
    for _ in range(nsys):
        t = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        mols.append(random_polyhex(rng, t))
    return mols


# ---------------------------------------------------------------------------
# surface
# ---------------------------------------------------------------------------
def rocksalt_slab(Lx: float, Ly: float, a: float = NACL_A, q: float = NACL_Q, nlayers: int = 2):
    """Rock-salt (001) slab filling an Lx x Ly periodic cell.

    Returns (elems, pos[N,3], charges[N]).  Top layer at z=0 with Na at
    (0,0); layers alternate every a/2 downwards with the species swapped."""
    nx = int(round(Lx / (a / 2)))
    ny = int(round(Ly / (a / 2)))
    if abs(nx * a / 2 - Lx) > 1e-9 or abs(ny * a / 2 - Ly) > 1e-9 or nx % 2 or ny % 2:
        raise ValueError(f"cell {Lx}x{Ly} is not commensurate with the rock-salt lattice a={a}")
    elems, pos, chg = [], [], []
    for il in range(nlayers):
        z = -il * a / 2
        for ix in range(nx):
            for iy in range(ny):
                na = ((ix + iy + il) % 2) == 0
                elems.append("Na" if na else "Cl")
                pos.append([ix * a / 2, iy * a / 2, z])
                chg.append(q if na else -q)
    return elems, np.array(pos, dtype=np.float64), np.array(chg, dtype=np.float64)


def place_on_surface(mol: Molecule, angle: float, xy, height: float) -> np.ndarray:
    """Rotate in-plane by ``angle`` about the centroid, move the centroid to
    (x, y, height)."""
    p = mol.apos - mol.apos.mean(axis=0)
    c, s = math.cos(angle), math.sin(angle)
    R = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    p = p @ R.T
    return p + np.array([xy[0], xy[1], height])


def uniform_batch_poses(nsys: int, L: float = 24.0, hmin: float = 3.0, hmax: float = 5.0,
                        seed: int = SEED_BASE + 1):
    """Per-replica (angle, x, y, height): rotation U(0,2pi), offset U(0,L),
    height U(hmin,hmax)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ang = rng.uniform(0.0, 2 * math.pi, nsys)
    xy = rng.uniform(0.0, L, (nsys, 2))
    h = rng.uniform(hmin, hmax, nsys)
    return ang, xy, h


def layer_lattice(nx: int = 16, ny: int = 32, spacing: float = 11.0, seed: int = SEED_BASE + 3,
                  hmin: float = 3.0, hmax: float = 4.0, jitter_deg: float = 2.0, origin=(5.5, 5.5)):
    """A monolayer of PTCDA-like molecules on an nx x ny square lattice.

    At 11 A spacing a single orientation overlaps neighbours (closest
    contacts 1.6 A), so the molecules alternate 0 / 90 degrees on the two
    checkerboard sublattices (herringbone-like; closest contacts 2.2 A), with
    a seeded rotation jitter of +-jitter_deg and heights U(hmin, hmax).
    Returns (molecule, [positions [38,3] per lattice site])."""
    mol = ptcda_like(SEED_BASE)
    rng = np.random.Generator(np.random.PCG64(seed))
    X = []
    for i in range(nx):
        for j in range(ny):
            th = (0.0 if (i + j) % 2 == 0 else 90.0) + rng.uniform(-jitter_deg, jitter_deg)
            h = rng.uniform(hmin, hmax)
            X.append(place_on_surface(mol, math.radians(th), (origin[0] + spacing * i, origin[1] + spacing * j), h))
    return mol, X


def poisson_arrivals(rate: float, n: int, seed: int):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate, n))


def lognormal_sizes(n: int, median: float = 60.0, sigma: float = 0.5, lo: int = 20, hi: int = 600, seed: int = 0):
    """Serving request sizes: lognormal (ln-sigma 0.5), clipped."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = np.exp(rng.normal(math.log(median), sigma, n))
    return np.clip(np.round(s), lo, hi).astype(int)


def raff_perturbed_replicas(pos, is_h, is_padding, nsys, perturb=0.2, seed=12345):
    """Replicas of a packed RAFF molecule with the random perturbations of the
    reference's RAFF benchmark (crates/libs/oclff/src/bin/
    raff_ocl_benchmark.rs:99-116): a 64-bit LCG (multiplier
    6364136223846793005, increment 1442695040888963407) draws, per replica,
    atom and coordinate, r = (state >> 33) / 2^31 - 0.5 and adds
    2 amp r with amp = perturb (1.5 perturb for H); padding atoms stay at 0.
    Returns float32 [nsys, natoms, 3]."""
    pos = np.asarray(pos, dtype=np.float32)
    n = len(pos)
    out = np.zeros((nsys, n, 3), dtype=np.float32)
    mask = (1 << 64) - 1
    s = seed
    for k in range(nsys):
        for i in range(n):
            if is_padding[i]:
                continue
            amp = np.float32(perturb * 1.5 if is_h[i] else perturb)
            for d in range(3):
                s = (s * 6364136223846793005 + 1442695040888963407) & mask
                r = np.float32(np.float32(s >> 33) / np.float32(1 << 31)) - np.float32(0.5)
                out[k, i, d] = pos[i, d] + amp * r * np.float32(2.0)
    return out


def xylitol_raff_molecule(golden_xyz):
    """Elements, float32 positions and bonds of xylitol as the reference's RAFF
    benchmark builds them (raff_ocl_benchmark.rs:24-47): a bond wherever
    d < r_i + r_j + 0.4 with covalent radii H 0.31, C 0.76, N 0.71, O 0.66."""
    rc = {"H": 0.31, "C": 0.76, "N": 0.71, "O": 0.66, "F": 0.57, "P": 1.07, "S": 1.05, "Cl": 1.02}
    el, x = [], []
    with open(golden_xyz) as f:
        lines = f.read().splitlines()
    n = int(lines[0].split()[0])
    for ln in lines[2:2 + n]:
        w = ln.split()
        el.append(w[0])
        x.append([float(w[1]), float(w[2]), float(w[3])])
    x = np.array(x)
    bonds = []
    for i in range(n):
        for j in range(i + 1, n):
            cut = rc.get(el[i], 0.7) + rc.get(el[j], 0.7) + 0.4
            if ((x[j] - x[i]) ** 2).sum() < cut * cut:
                bonds.append((i, j))
    return el, x.astype(np.float32), bonds
