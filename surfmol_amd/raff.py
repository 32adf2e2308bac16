"""Multi-replica rigid-atom port relaxation (RAFF "RRsp3" cluster step) over
the C ABI of include/surfmol_raff.h and the host set-up of surfmol_host.h.

* ``RaffMolecule.build``  -- one replica's tables: the set-up of the
                              reference's RAFF benchmark with oclff::pack
                              (crates/libs/oclff/src/pack.rs), in the
                              library's C++ host code.
* ``RaffReplicas``        -- oclff::RRsp3Multi (crates/libs/oclff/src/
                              rrsp3.rs:699-1088): shared topology, per-replica
                              state on the device, step / relax / downloads.
* ``RaffServer``          -- continuous-batching relax server for replicas of
                              one molecule.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib

_P = ctypes.POINTER


def _fp(a):
    return a.ctypes.data_as(_P(ctypes.c_float))


def _ip(a):
    return a.ctypes.data_as(_P(ctypes.c_int))


@dataclass
class StepConfig:
    """StepConfig (rrsp3.rs:54-73) with the reference's defaults."""
    dt: float = 0.1
    k_coll: float = 50.0
    relaxation: float = 0.5
    bbox_margin: float = 0.5
    momentum_beta: float = 0.0
    rot_mass_scale: float = 1.0
    n_rot_substeps: int = 5
    rot_eps: float = 0.0
    theta_max: float = 0.0
    damp: float = 1.0

    def cstruct(self):
        return _lib.SmrRaffStepConfig(self.dt, self.k_coll, self.relaxation, self.bbox_margin, self.momentum_beta,
                                      self.rot_mass_scale, self.n_rot_substeps, self.rot_eps, self.theta_max,
                                      self.damp)


class RaffMolecule:
    """One replica's packed tables (smh_raff_build): nodes first, caps, then
    padding to group_size (0: none)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        t = _lib.SmhRaffTables()
        self._check(lib().smh_raff_tables_of(self._h, ctypes.byref(t)))
        N, nr = t.natoms, t.nreal
        self.natoms, self.nreal, self.nnode = N, nr, t.nnode
        self.elems = t.elems.decode().split()
        self.pos3 = np.ctypeslib.as_array(t.pos3, (N, 3)).copy()
        self.inv_mass = np.ctypeslib.as_array(t.inv_mass, (N,)).copy()
        self.neighs = np.ctypeslib.as_array(t.neighs, (N, 4)).copy()
        self.excl1 = np.ctypeslib.as_array(t.excl1, (N, 4)).copy()
        self.excl2 = np.ctypeslib.as_array(t.excl2, (N, 4)).copy()
        self.bk_slots = np.ctypeslib.as_array(t.bk_slots, (N, 4)).copy()
        self.port_local = np.ctypeslib.as_array(t.port_local, (N, 4, 4)).copy()
        self.kflat = np.ctypeslib.as_array(t.kflat, (N, 4)).copy()
        self.is_padding = np.ctypeslib.as_array(t.is_padding, (N,)).astype(bool)
        self.perm = np.ctypeslib.as_array(t.perm, (nr,)).copy()

    @staticmethod
    def _check(rc):
        if rc != 0:
            raise _lib.SmrError(rc, lib().smh_raff_last_error().decode(errors="replace"))

    @classmethod
    def build(cls, elems, pos, bonds, nnode=-1, group_size=0, k_stiff=200.0):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = len(elems)
        arr = (ctypes.c_char_p * n)(*[e.encode() for e in elems])
        b = np.ascontiguousarray(bonds, dtype=np.int32).reshape(-1, 2)
        h = ctypes.c_void_p()
        cls._check(lib().smh_raff_build(n, arr, _fp(pos), len(b), _ip(b), int(nnode), int(group_size),
                                        float(k_stiff), ctypes.byref(h)))
        m = cls(h.value)
        return m

    def __del__(self):
        try:
            if self._h:
                lib().smh_raff_free(self._h)
                self._h = None
        except Exception:
            pass


class RaffReplicas:
    """nsys replicas of one molecule on the device (RRsp3Multi)."""

    def __init__(self, ctx, natoms_per_sys: int, nsys: int):
        self.ctx = ctx
        self.n, self.nsys = int(natoms_per_sys), int(nsys)
        self.nnode = None
        h = ctypes.c_void_p()
        check(lib().smr_raff_create(ctx._h, self.n, self.nsys, ctypes.byref(h)))
        self._h = h

    @classmethod
    def from_molecule(cls, ctx, mol: RaffMolecule, pos3_all, nsys, radius=None, fixmask=None, quat=None):
        """Upload the tables of `mol` (the benchmark's upload sequence,
        raff_ocl_benchmark.rs:125-129) and the replicas' positions
        [nsys, natoms, 3]."""
        r = cls(ctx, mol.natoms, nsys)
        r.upload_radius(np.zeros(mol.natoms, np.float32) if radius is None else radius)
        r.upload_neighs_and_exclusions(mol.neighs, mol.excl1, mol.excl2)
        r.upload_cluster_ports(mol.port_local, mol.kflat, mol.nnode)
        r.upload_bk_slots(mol.bk_slots)
        if fixmask is not None:
            r.upload_fixmask(fixmask)
        r.upload_state(pos3_all, mol.inv_mass, quat)
        return r

    def upload_radius(self, radius):
        a = np.ascontiguousarray(radius, dtype=np.float32)
        check(lib().smr_raff_upload_radius(self._h, _fp(a)))

    def upload_neighs_and_exclusions(self, neighs, excl1, excl2):
        a, b, c = (np.ascontiguousarray(x, dtype=np.int32) for x in (neighs, excl1, excl2))
        check(lib().smr_raff_upload_neighs_and_exclusions(self._h, _ip(a), _ip(b), _ip(c)))

    def upload_fixmask(self, fixmask):
        a = np.ascontiguousarray(fixmask, dtype=np.int32)
        check(lib().smr_raff_upload_fixmask(self._h, _ip(a)))

    def upload_cluster_ports(self, port_local_atoms, k_atoms, nnode):
        a = np.ascontiguousarray(port_local_atoms, dtype=np.float32)
        b = np.ascontiguousarray(k_atoms, dtype=np.float32)
        check(lib().smr_raff_upload_cluster_ports(self._h, _fp(a), _fp(b), int(nnode)))
        self.nnode = int(nnode)

    def upload_bk_slots(self, bk_slots):
        a = np.ascontiguousarray(bk_slots, dtype=np.int32)
        check(lib().smr_raff_upload_bk_slots(self._h, _ip(a)))

    def upload_state(self, pos3_all, inv_mass, quat=None):
        p = np.ascontiguousarray(pos3_all, dtype=np.float32)
        w = np.ascontiguousarray(inv_mass, dtype=np.float32)
        q = None if quat is None else np.ascontiguousarray(quat, dtype=np.float32)
        check(lib().smr_raff_upload_state(self._h, _fp(p), _fp(w), None if q is None else _fp(q)))

    def step(self, cfg: StepConfig, nsteps: int = 1):
        c = cfg.cstruct()
        check(lib().smr_raff_step(self._h, ctypes.byref(c), int(nsteps)))

    def relax(self, cfg: StepConfig, max_steps: int, tol: float):
        """Per-replica relaxation: (steps, status, metric) arrays [nsys]."""
        c = cfg.cstruct()
        st = np.zeros(self.nsys, np.int32)
        ss = np.zeros(self.nsys, np.int32)
        me = np.zeros(self.nsys, np.float32)
        check(lib().smr_raff_relax(self._h, ctypes.byref(c), int(max_steps), float(tol), _ip(st), _ip(ss), _fp(me)))
        return st, ss, me

    def positions(self):
        out = np.zeros((self.nsys, self.n, 3), np.float32)
        check(lib().smr_raff_download_pos(self._h, _fp(out)))
        return out

    def positions_replica(self, isys):
        out = np.zeros((self.n, 3), np.float32)
        check(lib().smr_raff_download_pos_replica(self._h, int(isys), _fp(out)))
        return out

    def quat_replica(self, isys):
        out = np.zeros((self.n, 4), np.float32)
        check(lib().smr_raff_download_quat_replica(self._h, int(isys), _fp(out)))
        return out

    def dpos_coll_replica(self, isys):
        out = np.zeros((self.n, 3), np.float32)
        check(lib().smr_raff_download_dpos_coll_replica(self._h, int(isys), _fp(out)))
        return out

    def dpos_node_replica(self, isys):
        out = np.zeros((max(1, self.nnode or 0), 3), np.float32)
        check(lib().smr_raff_download_dpos_node_replica(self._h, int(isys), _fp(out)))
        return out[:self.nnode]

    def metric(self):
        out = np.zeros(self.nsys, np.float32)
        check(lib().smr_raff_download_metric(self._h, _fp(out)))
        return out

    def close(self):
        if self._h:
            lib().smr_raff_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RaffServer:
    """smr_raff_server_*: submit replicas while others relax; each retires at
    convergence (tol) or at max_steps."""

    def __init__(self, template: RaffReplicas, capacity=1024, chunk_steps=64, max_steps=5000, tol=1e-4,
                 cfg: StepConfig | None = None):
        c = _lib.SmrRaffServerConfig()
        c.capacity, c.chunk_steps, c.max_steps, c.tol = int(capacity), int(chunk_steps), int(max_steps), float(tol)
        c.step = (cfg or StepConfig()).cstruct()
        h = ctypes.c_void_p()
        check(lib().smr_raff_server_create(template._h, ctypes.byref(c), ctypes.byref(h)))
        self._h = h
        self.n = template.n

    def submit(self, pos3, quat=None) -> int:
        p = np.ascontiguousarray(pos3, dtype=np.float32).reshape(self.n, 3)
        q = None if quat is None else np.ascontiguousarray(quat, dtype=np.float32).reshape(self.n, 4)
        rid = ctypes.c_longlong()
        check(lib().smr_raff_server_submit(self._h, _fp(p), None if q is None else _fp(q), ctypes.byref(rid)))
        return rid.value

    def poll(self, rid: int):
        """None while the request runs; else a dict (forgets the request)."""
        done, steps, status = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        met = ctypes.c_float()
        x = np.zeros((self.n, 3), np.float32)
        ts, td = ctypes.c_double(), ctypes.c_double()
        check(lib().smr_raff_server_poll(self._h, ctypes.c_longlong(rid), ctypes.byref(done), ctypes.byref(steps),
                                         ctypes.byref(status), ctypes.byref(met), _fp(x), ctypes.byref(ts),
                                         ctypes.byref(td)))
        if not done.value:
            return None
        return {"steps": steps.value, "status": status.value, "metric": met.value, "positions": x,
                "t_submit": ts.value, "t_done": td.value, "latency": td.value - ts.value}

    def drain(self):
        check(lib().smr_raff_server_drain(self._h))

    def stats(self):
        o = (ctypes.c_double * 7)()
        check(lib().smr_raff_server_stats(self._h, o))
        return dict(zip(["submitted", "retired", "resident", "queued", "chunks", "peak_resident", "device_s"], list(o)))

    def close(self):
        if self._h:
            lib().smr_raff_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
