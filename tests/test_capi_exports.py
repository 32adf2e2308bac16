"""The C-ABI library loads and exports every function declared in
include/*.h; the Python binding declares the same set.  No device calls."""
import os
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in ["surfmol_relax.h", "surfmol_host.h", "surfmol_raff.h"]:
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(smr_[a-z0-9_]+|smh_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    import surfmol_amd
    from surfmol_amd import _lib
    L = surfmol_amd.lib()
    decl = _declared()
    assert len(decl) > 30
    missing = [n for n in sorted(decl) if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTED_SYMBOLS) == decl


def test_version_and_error_channel():
    import surfmol_amd
    L = surfmol_amd.lib()
    assert b"gfx950" in L.smr_version()
    assert L.smr_ctx_create(0, None) == 1            # NULL out-pointer -> SMR_ERR_INVALID, no abort
    assert b"out == NULL" in L.smr_last_error()


def test_result_rows_match_the_c_struct():
    """Batch.relax_batched hands a numpy structured array to the library to be
    filled in place: its fields, offsets and size must be those of
    smr_relax_result as declared in include/surfmol_relax.h (C natural
    alignment)."""
    import numpy as np
    import surfmol_amd as S
    txt = open(os.path.join(ROOT, "include", "surfmol_relax.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    body = re.search(r"typedef struct \{([^}]*)\}\s*smr_relax_result;", txt).group(1)
    size = {"int": 4, "double": 8}
    off, fields = 0, []
    for m in re.finditer(r"(int|double)\s+([a-z_0-9]+)(?:\[(\d+)\])?\s*;", body):
        t, name, n = m.group(1), m.group(2), int(m.group(3) or 1)
        off = (off + size[t] - 1) // size[t] * size[t]
        fields.append((name, off, size[t] * n))
        off += size[t] * n
    off = (off + 7) // 8 * 8
    dt = S.RESULT_DTYPE
    assert list(dt.names) == [f[0] for f in fields]
    for name, o, nbytes in fields:
        assert dt.fields[name][1] == o, name
        assert dt.fields[name][0].itemsize == nbytes, name
    assert dt.itemsize == off
    rows = np.zeros(3, dtype=dt)
    rows[1]["steps"] = 7
    assert rows[1]["steps"] == 7 and rows.ctypes.data % 8 == 0
