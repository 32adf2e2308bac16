"""ctypes binding of libsurfmol_relax.so (include/surfmol_relax.h, surfmol_host.h,
surfmol_raff.h).

The library is built in-tree by ``surfmol_amd/build.py``.  Loading fails
loudly when it is missing: there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsurfmol_relax.so")
DATA_DIR = os.path.join(HERE, "data")

SMR_OK = 0
SMR_ERR_INVALID = 1
SMR_ERR_HIP = 2
SMR_ERR_NOMEM = 3
SMR_ERR_UNSUPPORTED = 4
SMR_ERR_STATE = 5
SMR_SYS_RUNNING = 0
SMR_SYS_CONVERGED = 1
SMR_SYS_MAXSTEPS = 2
SMR_SYS_NONFINITE = 3
SMR_RAFF_RUNNING = 0
SMR_RAFF_CONVERGED = 1
SMR_RAFF_MAXSTEPS = 2

P = ctypes.POINTER
c_int, c_double, c_float, c_char_p, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_char_p, ctypes.c_void_p


class SmrRaffStepConfig(ctypes.Structure):
    _fields_ = [("dt", c_float), ("k_coll", c_float), ("relaxation", c_float), ("bbox_margin", c_float),
                ("momentum_beta", c_float), ("rot_mass_scale", c_float), ("n_rot_substeps", c_int),
                ("rot_eps", c_float), ("theta_max", c_float), ("damp", c_float)]


class SmrRaffServerConfig(ctypes.Structure):
    _fields_ = [("capacity", c_int), ("chunk_steps", c_int), ("max_steps", c_int), ("tol", c_float),
                ("step", SmrRaffStepConfig)]


class SmhRaffTables(ctypes.Structure):
    _fields_ = [("natoms", c_int), ("nreal", c_int), ("nnode", c_int), ("pos3", P(c_float)), ("inv_mass", P(c_float)),
                ("neighs", P(c_int)), ("excl1", P(c_int)), ("excl2", P(c_int)), ("bk_slots", P(c_int)),
                ("port_local", P(c_float)), ("kflat", P(c_float)), ("is_padding", P(c_int)), ("perm", P(c_int)),
                ("elems", c_char_p)]


class SmrTopology(ctypes.Structure):
    _fields_ = [
        ("natoms", c_int), ("nbonds", c_int), ("nangles", c_int), ("ndih", c_int), ("ninv", c_int),
        ("neighs", P(c_int)), ("neigh_bs", P(c_int)), ("bon_params", P(c_double)),
        ("ang_atoms", P(c_int)), ("ang_ngs", P(c_int)), ("ang_params", P(c_double)),
        ("dih_atoms", P(c_int)), ("dih_ngs", P(c_int)), ("dih_params", P(c_double)),
        ("inv_atoms", P(c_int)), ("inv_ngs", P(c_int)), ("inv_params", P(c_double)),
        ("reqs", P(c_double)),
    ]


class SmrGridDesc(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("p0", c_double * 3), ("step", c_double * 3),
                ("alpha_morse", c_double), ("rdamp", c_double), ("keep_reference", c_int)]


class SmrFFOptions(ctypes.Structure):
    _fields_ = [("rdamp", c_double), ("nonbonded", c_int), ("surface", c_int), ("inter_cutoff", c_double),
                ("inter_switch", c_double), ("pbc_lx", c_double), ("pbc_ly", c_double)]


class SmrFireParams(ctypes.Structure):
    _fields_ = [("dt0", c_double), ("dt_max", c_double), ("alpha0", c_double), ("n_min", c_int),
                ("f_inc", c_double), ("f_dec", c_double), ("f_alpha", c_double), ("flim", c_double)]


class SmrRelaxResult(ctypes.Structure):
    _fields_ = [("steps", c_int), ("status", c_int), ("energy", c_double), ("fmax", c_double),
                ("terms", c_double * 6)]


class SmrServerConfig(ctypes.Structure):
    _fields_ = [("max_resident", c_int), ("max_atoms", c_int), ("hbm_fraction", c_double),
                ("chunk_steps", c_int), ("max_steps", c_int), ("fconv", c_double),
                ("fire", SmrFireParams), ("ff", SmrFFOptions)]


# (name, restype, argtypes) of every exported entry point
_SIGNATURES = [
    ("smr_last_error", c_char_p, []),
    ("smr_version", c_char_p, []),
    ("smr_ctx_create", c_int, [c_int, P(c_void_p)]),
    ("smr_ctx_destroy", c_int, [c_void_p]),
    ("smr_ctx_info", c_int, [c_void_p, P(c_double)]),
    ("smr_gridff_upload", c_int, [c_void_p, P(SmrGridDesc), P(c_float), P(c_void_p)]),
    ("smr_gridff_build", c_int, [c_void_p, P(SmrGridDesc), c_int, P(c_double), P(c_double), c_int, P(c_void_p)]),
    ("smr_gridff_download", c_int, [c_void_p, P(c_float)]),
    ("smr_gridff_download_reference", c_int, [c_void_p, P(c_double)]),
    ("smr_gridff_destroy", c_int, [c_void_p]),
    ("smr_gridff_save", c_int, [c_void_p, c_char_p]),
    ("smr_gridff_load", c_int, [c_void_p, P(SmrGridDesc), c_char_p, P(c_void_p)]),
    ("smr_gridff_build_cached", c_int, [c_void_p, P(SmrGridDesc), c_int, P(c_double), P(c_double), c_int, c_char_p,
                                        P(c_int), P(c_void_p)]),
    ("smr_batch_create", c_int, [c_void_p, c_void_p, c_int, P(SmrTopology), c_int, P(c_int), P(c_double),
                                 P(SmrFFOptions), P(c_void_p)]),
    ("smr_batch_destroy", c_int, [c_void_p]),
    ("smr_batch_natoms", c_int, [c_void_p, P(c_int)]),
    ("smr_batch_atom_offsets", c_int, [c_void_p, P(c_int)]),
    ("smr_batch_set_positions", c_int, [c_void_p, P(c_double)]),
    ("smr_batch_get_positions", c_int, [c_void_p, P(c_double)]),
    ("smr_batch_reset", c_int, [c_void_p]),
    ("smr_batch_eval_forces", c_int, [c_void_p, P(c_double), P(c_double)]),
    ("smr_relax_batched", c_int, [c_void_p, P(SmrFireParams), c_int, c_double, P(SmrRelaxResult)]),
    ("smr_relax_plain", c_int, [c_void_p, c_int, P(SmrFireParams), c_int, c_double, P(SmrRelaxResult)]),
    ("smr_relax_plain_noalloc", c_int, [c_void_p, c_int, P(SmrFireParams), c_int, c_double, P(SmrRelaxResult)]),
    ("smr_run_md", c_int, [c_void_p, c_int, c_double, c_double, c_double, c_double, P(c_int)]),
    ("smr_batch_last_launch", c_int, [c_void_p, P(c_int), P(c_float)]),
    ("smr_batch_set_launch", c_int, [c_void_p, c_int, c_int, c_int]),
    ("smr_batch_plan", c_int, [c_void_p, P(c_int), P(c_int)]),
    ("smr_batch_autotune", c_int, [c_void_p, P(SmrFireParams), c_int]),
    ("smr_batch_set_graph", c_int, [c_void_p, c_int]),
    ("smr_batch_set_layer_persistent", c_int, [c_void_p, c_int, c_double]),
    ("smr_batch_layer_path", c_int, [c_void_p, P(c_int), P(c_int), P(c_int)]),
    ("smr_batch_set_order", c_int, [c_void_p, c_int]),
    ("smr_batch_set_rerank", c_int, [c_void_p, c_int]),
    ("smr_server_create", c_int, [c_void_p, c_void_p, P(SmrServerConfig), P(c_void_p)]),
    ("smr_server_submit", c_int, [c_void_p, P(SmrTopology), P(c_double), P(ctypes.c_longlong)]),
    ("smr_server_poll", c_int, [c_void_p, ctypes.c_longlong, P(c_int), P(SmrRelaxResult), P(c_double),
                                P(c_double), P(c_double)]),
    ("smr_server_drain", c_int, [c_void_p]),
    ("smr_server_stats", c_int, [c_void_p, P(c_double)]),
    ("smr_server_destroy", c_int, [c_void_p]),
    ("smh_last_error", c_char_p, []),
    ("smh_molecule_build", c_int, [c_char_p, c_int, P(c_char_p), P(c_double), P(c_double), c_int, P(c_int),
                                   c_double, P(c_void_p)]),
    ("smh_molecule_counts", c_int, [c_void_p, P(c_int)]),
    ("smh_molecule_topology", c_int, [c_void_p, P(SmrTopology)]),
    ("smh_molecule_bonds", c_int, [c_void_p, P(c_int)]),
    ("smh_molecule_types", c_char_p, [c_void_p]),
    ("smh_molecule_free", None, [c_void_p]),
    ("smh_molecule_read_ufftopo", c_int, [c_char_p, c_char_p, P(c_double), P(c_void_p)]),
    ("smh_molecule_read_json", c_int, [c_char_p, c_char_p, P(c_double), P(c_void_p)]),
    ("smh_molecule_write_ufftopo", c_int, [c_void_p, c_char_p]),
    ("smh_molecule_write_json", c_int, [c_void_p, c_char_p]),
    ("smh_molecule_positions", c_int, [c_void_p, P(c_double)]),
    ("smh_raff_last_error", c_char_p, []),
    ("smh_raff_build", c_int, [c_int, P(c_char_p), P(c_float), c_int, P(c_int), c_int, c_int, c_float, P(c_void_p)]),
    ("smh_raff_tables_of", c_int, [c_void_p, P(SmhRaffTables)]),
    ("smh_raff_free", None, [c_void_p]),
    ("smr_raff_create", c_int, [c_void_p, c_int, c_int, P(c_void_p)]),
    ("smr_raff_destroy", c_int, [c_void_p]),
    ("smr_raff_upload_radius", c_int, [c_void_p, P(c_float)]),
    ("smr_raff_upload_neighs_and_exclusions", c_int, [c_void_p, P(c_int), P(c_int), P(c_int)]),
    ("smr_raff_upload_fixmask", c_int, [c_void_p, P(c_int)]),
    ("smr_raff_upload_cluster_ports", c_int, [c_void_p, P(c_float), P(c_float), c_int]),
    ("smr_raff_upload_bk_slots", c_int, [c_void_p, P(c_int)]),
    ("smr_raff_upload_state", c_int, [c_void_p, P(c_float), P(c_float), P(c_float)]),
    ("smr_raff_step", c_int, [c_void_p, P(SmrRaffStepConfig), c_int]),
    ("smr_raff_relax", c_int, [c_void_p, P(SmrRaffStepConfig), c_int, c_float, P(c_int), P(c_int), P(c_float)]),
    ("smr_raff_download_pos", c_int, [c_void_p, P(c_float)]),
    ("smr_raff_download_pos_replica", c_int, [c_void_p, c_int, P(c_float)]),
    ("smr_raff_download_quat_replica", c_int, [c_void_p, c_int, P(c_float)]),
    ("smr_raff_download_dpos_coll_replica", c_int, [c_void_p, c_int, P(c_float)]),
    ("smr_raff_download_dpos_node_replica", c_int, [c_void_p, c_int, P(c_float)]),
    ("smr_raff_download_metric", c_int, [c_void_p, P(c_float)]),
    ("smr_raff_server_create", c_int, [c_void_p, P(SmrRaffServerConfig), P(c_void_p)]),
    ("smr_raff_server_submit", c_int, [c_void_p, P(c_float), P(c_float), P(ctypes.c_longlong)]),
    ("smr_raff_server_poll", c_int, [c_void_p, ctypes.c_longlong, P(c_int), P(c_int), P(c_int), P(c_float), P(c_float),
                                     P(c_double), P(c_double)]),
    ("smr_raff_server_drain", c_int, [c_void_p]),
    ("smr_raff_server_stats", c_int, [c_void_p, P(c_double)]),
    ("smr_raff_server_destroy", c_int, [c_void_p]),
]

EXPORTED_SYMBOLS = [s[0] for s in _SIGNATURES]

_lib = None


class SmrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[smr code {code}] {msg}")
        self.code = code


def lib():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python surfmol_amd/build.py` "
                               f"(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, host=False):
    if rc != 0:
        L = lib()
        msg = (L.smh_last_error() if host else L.smr_last_error()).decode(errors="replace")
        raise SmrError(rc, msg)
    return rc
