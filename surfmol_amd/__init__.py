"""surfmol_amd -- MI355X (gfx950) relaxation path for SurfMol systems.

Python host mirror of the reference's set-up + relax API over the C ABI of
``libsurfmol_relax.so`` (include/surfmol_relax.h, include/surfmol_host.h):

* ``Molecule.build``   -- moltopo/molff set-up (topology.rs, assign_uff.rs,
                          params.rs, uff.rs:699-887) in the library's C++ host
                          code; yields the Uff/NonBondedFF tables.
* ``Context``          -- one device + stream (oclff::nvidia_proque analogue).
* ``GridFF``           -- tricubic B-spline surface grid: device FP64 build +
                          fit, or upload of fitted coefficients.
* ``Batch``            -- many systems sharing topologies: eval_forces
                          (MolWorld::eval_forces), relax_batched (FIRE,
                          persistent kernel), relax_plain (one launch sequence
                          per step, host FIRE), run_md (MolWorld::run_md).
                          With inter_cutoff > 0 the systems form one coupled
                          layer (intermolecular pairs through a cell list).
* ``Server``           -- continuous-batching relax server.

All compute runs in HIP kernels; importing works without a GPU, creating a
Context needs one.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (DATA_DIR, SMR_SYS_CONVERGED, SMR_SYS_MAXSTEPS, SMR_SYS_NONFINITE, SMR_SYS_RUNNING, SmrError,
                   check, lib)

__all__ = ["Molecule", "Context", "GridFF", "GridSpec", "Batch", "Server", "FireParams", "substrate_reqs", "SmrError",
           "RESULT_DTYPE", "to_dicts",
           "DATA_DIR", "SMR_SYS_CONVERGED",
           "SMR_SYS_MAXSTEPS", "SMR_SYS_NONFINITE", "SMR_SYS_RUNNING", "lib"]

_P = ctypes.POINTER


def _dptr(a):
    return a.ctypes.data_as(_P(ctypes.c_double))


def _iptr(a):
    return a.ctypes.data_as(_P(ctypes.c_int))


class Molecule:
    """One molecule's force-field tables, built by the library's C++ host
    code (the reference's set-up sequence, relax_pentacene_uff.rs:53-78)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        self.topo = _lib.SmrTopology()
        check(lib().smh_molecule_topology(self._h, ctypes.byref(self.topo)), host=True)

    @classmethod
    def build(cls, elems, apos, charges=None, bonds=None, rcut_bond=1.8, param_dir=DATA_DIR):
        apos = np.ascontiguousarray(apos, dtype=np.float64).reshape(-1, 3)
        n = len(elems)
        if apos.shape[0] != n:
            raise ValueError(f"{n} elements but {apos.shape[0]} positions")
        arr = (ctypes.c_char_p * n)(*[e.encode() for e in elems])
        q = None if charges is None else np.ascontiguousarray(charges, dtype=np.float64)
        if bonds is None:
            nb, bp = 0, None
        else:
            b = np.ascontiguousarray(bonds, dtype=np.int32).reshape(-1, 2)
            nb, bp = len(b), b
        h = ctypes.c_void_p()
        check(lib().smh_molecule_build(param_dir.encode(), n, arr, _dptr(apos), None if q is None else _dptr(q), nb,
                                       None if bp is None else _iptr(bp), float(rcut_bond), ctypes.byref(h)), host=True)
        m = cls(h.value)
        m._keep = (arr, apos, q, bp)
        return m

    @classmethod
    def _read(cls, fn, path, charges, param_dir):
        q = None if charges is None else np.ascontiguousarray(charges, dtype=np.float64)
        h = ctypes.c_void_p()
        check(fn(param_dir.encode(), os.fspath(path).encode(), None if q is None else _dptr(q), ctypes.byref(h)),
              host=True)
        m = cls(h.value)
        if q is not None and len(q) != m.natoms:
            raise ValueError(f"{len(q)} charges for {m.natoms} atoms")
        m._keep = (q,)
        return m

    @classmethod
    def read_ufftopo(cls, path, charges=None, param_dir=DATA_DIR):
        """A topology from buildff's UFFTOPO binary (crates/apps/buildff/src/
        main.rs:10-30): its term lists and UFF types as given."""
        return cls._read(lib().smh_molecule_read_ufftopo, path, charges, param_dir)

    @classmethod
    def read_json(cls, path, charges=None, param_dir=DATA_DIR):
        """A topology from a TopologyData JSON (moltopo/src/export.rs:12-26)."""
        return cls._read(lib().smh_molecule_read_json, path, charges, param_dir)

    def write_ufftopo(self, path):
        check(lib().smh_molecule_write_ufftopo(self._h, os.fspath(path).encode()), host=True)

    def write_json(self, path):
        check(lib().smh_molecule_write_json(self._h, os.fspath(path).encode()), host=True)

    def positions(self):
        x = np.zeros((self.natoms, 3))
        check(lib().smh_molecule_positions(self._h, _dptr(x)), host=True)
        return x

    def __del__(self):
        try:
            if self._h:
                lib().smh_molecule_free(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def counts(self):
        c = (ctypes.c_int * 5)()
        check(lib().smh_molecule_counts(self._h, c), host=True)
        return tuple(c)

    @property
    def natoms(self):
        return self.topo.natoms

    @property
    def types(self):
        return lib().smh_molecule_types(self._h).decode().split()

    def tables(self):
        """numpy copies of every table (for checking against the oracle)."""
        t = self.topo
        n, nb, na, nd, ni = self.counts

        def arr(p, k, dt):
            if k == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(p, shape=(k,)).copy()

        bonds = np.zeros(2 * nb, dtype=np.int32)
        if nb:
            check(lib().smh_molecule_bonds(self._h, _iptr(bonds)), host=True)
        return {
            "bonds": bonds.reshape(-1, 2),
            "neighs": arr(t.neighs, 4 * n, np.int32).reshape(-1, 4),
            "neigh_bs": arr(t.neigh_bs, 4 * n, np.int32).reshape(-1, 4),
            "bon_params": arr(t.bon_params, 2 * nb, np.float64).reshape(-1, 2),
            "ang_atoms": arr(t.ang_atoms, 3 * na, np.int32).reshape(-1, 3),
            "ang_ngs": arr(t.ang_ngs, 2 * na, np.int32).reshape(-1, 2),
            "ang_params": arr(t.ang_params, 5 * na, np.float64).reshape(-1, 5),
            "dih_atoms": arr(t.dih_atoms, 4 * nd, np.int32).reshape(-1, 4),
            "dih_ngs": arr(t.dih_ngs, 3 * nd, np.int32).reshape(-1, 3),
            "dih_params": arr(t.dih_params, 3 * nd, np.float64).reshape(-1, 3),
            "inv_atoms": arr(t.inv_atoms, 4 * ni, np.int32).reshape(-1, 4),
            "inv_ngs": arr(t.inv_ngs, 3 * ni, np.int32).reshape(-1, 3),
            "inv_params": arr(t.inv_params, 4 * ni, np.float64).reshape(-1, 4),
            "reqs": arr(t.reqs, 4 * n, np.float64).reshape(-1, 4),
            "types": self.types,
        }


def substrate_reqs(elems, charges, param_dir=DATA_DIR):
    """REQH rows (RvdW, sqrt(EvdW), Q, Hb) of isolated substrate atoms
    (params.rs:595-604 lookup by element type), with the given charges."""
    pos = np.zeros((len(elems), 3))
    pos[:, 0] = np.arange(len(elems)) * 10.0
    m = Molecule.build(list(elems), pos, charges, rcut_bond=0.0, param_dir=param_dir)
    r = m.tables()["reqs"]
    r[:, 3] = 0.0
    return r


class Context:
    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().smr_ctx_create(int(device), ctypes.byref(h)))
        self._h = h

    def info(self):
        o = (ctypes.c_double * 5)()
        check(lib().smr_ctx_info(self._h, o))
        return {"free_bytes": o[0], "total_bytes": o[1], "compute_units": int(o[2]), "lds_per_workgroup": int(o[3]),
                "wavefront": int(o[4])}

    def close(self):
        if self._h:
            lib().smr_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class GridSpec:
    nx: int
    ny: int
    nz: int
    p0: tuple = (0.0, 0.0, 0.0)
    step: tuple = (0.1, 0.1, 0.1)
    alpha_morse: float = 1.5
    rdamp: float = 1.0

    def desc(self, keep_reference=False):
        d = _lib.SmrGridDesc()
        d.nx, d.ny, d.nz = int(self.nx), int(self.ny), int(self.nz)
        for k in range(3):
            d.p0[k] = float(self.p0[k])
            d.step[k] = float(self.step[k])
        d.alpha_morse = float(self.alpha_morse)
        d.rdamp = float(self.rdamp)
        d.keep_reference = 1 if keep_reference else 0
        return d


class GridFF:
    def __init__(self, ctx: Context, handle, spec: GridSpec):
        self.ctx = ctx
        self._h = handle
        self.spec = spec

    @classmethod
    def build(cls, ctx: Context, spec: GridSpec, sub_pos, sub_reqs, npbc=2, keep_reference=False):
        sp = np.ascontiguousarray(sub_pos, dtype=np.float64).reshape(-1, 3)
        sr = np.ascontiguousarray(sub_reqs, dtype=np.float64).reshape(-1, 4)
        h = ctypes.c_void_p()
        d = spec.desc(keep_reference)
        check(lib().smr_gridff_build(ctx._h, ctypes.byref(d), len(sp), _dptr(sp), _dptr(sr), int(npbc), ctypes.byref(h)))
        return cls(ctx, h, spec)

    @classmethod
    def build_cached(cls, ctx: Context, spec: GridSpec, sub_pos, sub_reqs, cache_dir, npbc=2):
        """Load the fitted grid of these inputs from cache_dir, or build it and
        write it there.  Returns (grid, hit)."""
        sp = np.ascontiguousarray(sub_pos, dtype=np.float64).reshape(-1, 3)
        sr = np.ascontiguousarray(sub_reqs, dtype=np.float64).reshape(-1, 4)
        h = ctypes.c_void_p()
        hit = ctypes.c_int(0)
        d = spec.desc()
        check(lib().smr_gridff_build_cached(ctx._h, ctypes.byref(d), len(sp), _dptr(sp), _dptr(sr), int(npbc),
                                            os.fspath(cache_dir).encode(), ctypes.byref(hit), ctypes.byref(h)))
        return cls(ctx, h, spec), bool(hit.value)

    @classmethod
    def load(cls, ctx: Context, spec: GridSpec, path):
        """A grid from a coefficient file written by save() (.npy, float32,
        shape (nx, ny, nz, 4))."""
        h = ctypes.c_void_p()
        d = spec.desc()
        check(lib().smr_gridff_load(ctx._h, ctypes.byref(d), os.fspath(path).encode(), ctypes.byref(h)))
        return cls(ctx, h, spec)

    def save(self, path):
        check(lib().smr_gridff_save(self._h, os.fspath(path).encode()))

    @classmethod
    def upload(cls, ctx: Context, spec: GridSpec, coefs):
        c = np.ascontiguousarray(coefs, dtype=np.float32).reshape(-1)
        if c.size != spec.nx * spec.ny * spec.nz * 4:
            raise ValueError("coefficient array size does not match the grid")
        h = ctypes.c_void_p()
        d = spec.desc()
        check(lib().smr_gridff_upload(ctx._h, ctypes.byref(d), c.ctypes.data_as(_P(ctypes.c_float)), ctypes.byref(h)))
        return cls(ctx, h, spec)

    def coefficients(self):
        s = self.spec
        out = np.zeros((s.nx, s.ny, s.nz, 4), dtype=np.float32)
        check(lib().smr_gridff_download(self._h, out.ctypes.data_as(_P(ctypes.c_float))))
        return out

    def reference_values(self):
        s = self.spec
        out = np.zeros((s.nx, s.ny, s.nz, 4), dtype=np.float64)
        check(lib().smr_gridff_download_reference(self._h, _dptr(out)))
        return out

    def close(self):
        if self._h:
            lib().smr_gridff_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class FireParams:
    """FireState (molff/src/raff.rs:1040-1062); times in internal units."""
    dt0: float
    dt_max: float
    alpha0: float = 0.1
    n_min: int = 5
    f_inc: float = 1.1
    f_dec: float = 0.5
    f_alpha: float = 0.99
    flim: float = 0.0

    def cstruct(self):
        return _lib.SmrFireParams(self.dt0, self.dt_max, self.alpha0, self.n_min, self.f_inc, self.f_dec, self.f_alpha,
                                  self.flim)


# rows of smr_relax_result as numpy sees them (the ctypes layout)
RESULT_DTYPE = np.dtype(_lib.SmrRelaxResult)
assert RESULT_DTYPE.itemsize == ctypes.sizeof(_lib.SmrRelaxResult)


def _results(arr):
    return [{"steps": r.steps, "status": r.status, "energy": r.energy, "fmax": r.fmax, "terms": list(r.terms)}
            for r in arr]


def to_dicts(results):
    """Relax results (a RESULT_DTYPE array or row, as relax_batched and
    relax_plain return) as a list of plain Python dicts -- int/float fields and
    a list of six energy terms -- e.g. for json.dumps."""
    rows = np.atleast_1d(results)
    return [{"steps": int(r["steps"]), "status": int(r["status"]), "energy": float(r["energy"]),
             "fmax": float(r["fmax"]), "terms": [float(x) for x in r["terms"]]} for r in rows]


class Batch:
    """A set of systems on the device (shared topologies, per-system state)."""

    def __init__(self, ctx: Context, molecules, sys_topo, apos, grid: GridFF | None = None, rdamp=1.0,
                 nonbonded=True, surface=True, inter_cutoff=0.0, inter_switch=0.0, pbc=(0.0, 0.0)):
        self.ctx = ctx
        self.grid = grid
        self.molecules = list(molecules)
        topos = (_lib.SmrTopology * len(self.molecules))(*[m.topo for m in self.molecules])
        st = np.ascontiguousarray(sys_topo, dtype=np.int32)
        x = np.ascontiguousarray(apos, dtype=np.float64).reshape(-1, 3)
        nat = sum(self.molecules[t].natoms for t in st)
        if x.shape[0] != nat:
            raise ValueError(f"apos has {x.shape[0]} atoms, systems need {nat}")
        opts = _lib.SmrFFOptions(float(rdamp), 1 if nonbonded else 0, 1 if (surface and grid is not None) else 0,
                                 float(inter_cutoff), float(inter_switch), float(pbc[0]), float(pbc[1]))
        h = ctypes.c_void_p()
        check(lib().smr_batch_create(ctx._h, grid._h if grid is not None else None, len(self.molecules), topos,
                                     len(st), _iptr(st), _dptr(x), ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        self.nsys = len(st)
        self.natoms = nat
        off = np.zeros(self.nsys + 1, dtype=np.int32)
        check(lib().smr_batch_atom_offsets(self._h, _iptr(off)))
        self.offsets = off

    def positions(self):
        x = np.zeros((self.natoms, 3))
        check(lib().smr_batch_get_positions(self._h, _dptr(x)))
        return x

    def set_positions(self, apos):
        x = np.ascontiguousarray(apos, dtype=np.float64).reshape(-1, 3)
        check(lib().smr_batch_set_positions(self._h, _dptr(x)))

    def reset(self):
        """Back to the initial positions, zero velocities (device-side)."""
        check(lib().smr_batch_reset(self._h))

    def eval_forces(self):
        """(F[natoms,3], E[nsys,6]) at the current positions."""
        f = np.zeros((self.natoms, 3))
        e = np.zeros((self.nsys, 6))
        check(lib().smr_batch_eval_forces(self._h, _dptr(f), _dptr(e)))
        return f, e

    def relax_batched(self, fire: FireParams, max_steps: int, fconv: float):
        """FIRE on every system.  Returns a numpy structured array, one row per
        system, with the fields of smr_relax_result (steps, status, energy,
        fmax, terms): the library writes the rows in place, so a large batch
        pays no per-system conversion (row["steps"] reads like a dict)."""
        res = np.zeros(self.nsys, dtype=RESULT_DTYPE)
        fp = fire.cstruct()
        check(lib().smr_relax_batched(self._h, ctypes.byref(fp), int(max_steps), float(fconv),
                                      res.ctypes.data_as(ctypes.POINTER(_lib.SmrRelaxResult))))
        return res

    def relax_plain(self, isys: int, fire: FireParams, max_steps: int, fconv: float):
        """The plain method on system isys.  Returns one row of the type
        relax_batched returns (RESULT_DTYPE; ``to_dicts`` converts)."""
        res = np.zeros(1, dtype=RESULT_DTYPE)
        fp = fire.cstruct()
        check(lib().smr_relax_plain(self._h, int(isys), ctypes.byref(fp), int(max_steps), float(fconv),
                                    res.ctypes.data_as(ctypes.POINTER(_lib.SmrRelaxResult))))
        return res[0]

    def run_md(self, niter, dt, fconv, flim, damping):
        steps = np.zeros(self.nsys, dtype=np.int32)
        check(lib().smr_run_md(self._h, int(niter), float(dt), float(fconv), float(flim), float(damping), _iptr(steps)))
        return steps

    def set_launch(self, block=0, stencil_cache=-1, pair_sweep=-1):
        """Kernel variant (threads/workgroup, surface stencil cache, unordered-pair
        sweep); 0/-1/-1 = automatic."""
        check(lib().smr_batch_set_launch(self._h, int(block), int(stencil_cache), int(pair_sweep)))

    def last_launch(self):
        shape = (ctypes.c_int * 4)()
        ms = ctypes.c_float()
        check(lib().smr_batch_last_launch(self._h, shape, ctypes.byref(ms)))
        return {"block": shape[0], "apt": shape[1], "grid": shape[2], "lds_bytes": shape[3] & ((1 << 30) - 1),
                "stencil_cache": bool(shape[3] >> 30), "kernel_ms": ms.value}

    def plan(self):
        """Persistent launches of the batch: one dict per launch with its shape
        (adjacent size classes of one shape share a launch; "cls" is the
        largest class in it)."""
        n = ctypes.c_int()
        info = (ctypes.c_int * (8 * 8))()
        check(lib().smr_batch_plan(self._h, ctypes.byref(n), info))
        out = []
        for q in range(n.value):
            o = info[8 * q:8 * q + 8]
            out.append({"cls": o[0], "systems": o[1], "max_atoms": o[2], "block": o[3], "apt": o[4], "grid": o[5],
                        "lds_bytes": o[6], "fint_lds": bool(o[7] & 1), "pair_sweep": bool(o[7] & 2),
                        "stencil_cache": bool(o[7] & 4)})
        return out

    def autotune(self, fire: FireParams, trial_steps: int = 50):
        """Time the launch shapes of every size class; the context keeps the fastest."""
        fp = fire.cstruct()
        check(lib().smr_batch_autotune(self._h, ctypes.byref(fp), int(trial_steps)))
        return self.plan()

    def set_order(self, energy_order: bool):
        """True (default): start the systems of highest initial energy per
        atom first in each size class; False: batch order."""
        check(lib().smr_batch_set_order(self._h, 1 if energy_order else 0))

    def set_rerank(self, steps: int):
        """Re-rank the unfinished systems by the energy of their current
        positions after `steps` FIRE steps and continue in that order (two
        launches); -1 (default) = after max_steps / 5; 0 = one launch.
        Results do not depend on it."""
        check(lib().smr_batch_set_rerank(self._h, int(steps)))

    def set_graph(self, use_graph: bool):
        check(lib().smr_batch_set_graph(self._h, 1 if use_graph else 0))

    LAYER_FALLBACK = {0: None, 1: "grid barrier timed out", 2: "neighbour instances overflowed",
                      3: "staged neighbour atoms overflowed", 4: "neighbour list overflowed",
                      5: "the layer does not fit the chip at once", 6: "molecules need another launch shape"}

    def set_layer_persistent(self, on: bool = True, skin: float = 0.0):
        """Coupled layers: run the relaxation as one cooperative launch where
        the layer fits the chip (default), with Verlet lists of radius
        cutoff + skin (A; 0 keeps the current skin, 1 A at first); False uses
        the step launches.  Results agree to FP32 summation order."""
        check(lib().smr_batch_set_layer_persistent(self._h, 1 if on else 0, float(skin)))

    def layer_path(self) -> dict:
        """Which path the last coupled-layer relaxation took: persistent
        (one launch), the reason it fell back to the step launches, and the
        number of neighbour-list builds."""
        p, f, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().smr_batch_layer_path(self._h, ctypes.byref(p), ctypes.byref(f), ctypes.byref(r)))
        return {"persistent": bool(p.value), "fallback": self.LAYER_FALLBACK.get(f.value, f.value),
                "rebuilds": r.value}

    def close(self):
        if self._h:
            lib().smr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching relax server (smr_server_*): submit requests of any
    size while others relax; each retires at convergence."""

    def __init__(self, ctx: Context, grid: GridFF | None, fire: FireParams, max_resident=1024, max_atoms=600,
                 hbm_fraction=0.25, chunk_steps=64, max_steps=2000, fconv=1e-3, rdamp=1.0, nonbonded=True,
                 surface=True):
        self.ctx = ctx
        self.grid = grid
        cfg = _lib.SmrServerConfig()
        cfg.max_resident = int(max_resident)
        cfg.max_atoms = int(max_atoms)
        cfg.hbm_fraction = float(hbm_fraction)
        cfg.chunk_steps = int(chunk_steps)
        cfg.max_steps = int(max_steps)
        cfg.fconv = float(fconv)
        cfg.fire = fire.cstruct()
        cfg.ff = _lib.SmrFFOptions(float(rdamp), 1 if nonbonded else 0, 1 if (surface and grid is not None) else 0,
                                   0.0, 0.0)
        h = ctypes.c_void_p()
        check(lib().smr_server_create(ctx._h, grid._h if grid is not None else None, ctypes.byref(cfg),
                                      ctypes.byref(h)))
        self._h = h
        self._natoms = {}

    def submit(self, molecule: Molecule, apos) -> int:
        x = np.ascontiguousarray(apos, dtype=np.float64).reshape(-1, 3)
        if x.shape[0] != molecule.natoms:
            raise ValueError(f"{x.shape[0]} positions for {molecule.natoms} atoms")
        rid = ctypes.c_longlong()
        check(lib().smr_server_submit(self._h, ctypes.byref(molecule.topo), _dptr(x), ctypes.byref(rid)))
        self._natoms[rid.value] = molecule.natoms
        return rid.value

    def poll(self, rid: int, positions=True):
        """None while the request runs; else a dict (forgets the request) with
        plain Python values: steps, status, energy, fmax, terms (list), the
        submit/retire times and latency [s], and positions [N,3] when asked."""
        done = ctypes.c_int()
        res = _lib.SmrRelaxResult()
        n = self._natoms.get(rid)
        positions = positions and n is not None
        x = np.zeros((n, 3)) if positions else None
        ts, td = ctypes.c_double(), ctypes.c_double()
        check(lib().smr_server_poll(self._h, ctypes.c_longlong(rid), ctypes.byref(done), ctypes.byref(res),
                                    _dptr(x) if positions else None, ctypes.byref(ts), ctypes.byref(td)))
        if not done.value:
            return None
        self._natoms.pop(rid, None)
        out = _results([res])[0]
        out.update(t_submit=ts.value, t_done=td.value, latency=td.value - ts.value)
        if positions:
            out["positions"] = x
        return out

    def drain(self):
        check(lib().smr_server_drain(self._h))

    def stats(self):
        o = (ctypes.c_double * 10)()
        check(lib().smr_server_stats(self._h, o))
        keys = ["submitted", "retired", "resident", "queued", "chunks", "peak_resident", "arena_bytes",
                "arena_bytes_used", "atom_capacity", "device_s"]
        return dict(zip(keys, list(o)))

    def close(self):
        if self._h:
            lib().smr_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
