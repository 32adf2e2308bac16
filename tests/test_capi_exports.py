"""The C-ABI library loads and exports every function declared in
include/*.h; the Python binding declares the same set.  No device calls."""
import os
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in ["surfmol_relax.h", "surfmol_host.h"]:
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
