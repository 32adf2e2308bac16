"""Shared fixtures.  GPU tests are marked ``gpu`` and run on an MI355X; the
rest run on the CPU (oracle pins, host set-up, library exports)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def params(oracle):
    import surfmol_amd
    return oracle.Params(surfmol_amd.DATA_DIR)


@pytest.fixture(scope="session")
def ctx():
    import surfmol_amd as S
    c = S.Context(0)
    yield c
    # grids/batches still referenced (e.g. by a failing test's traceback) keep
    # the device context alive; close() only drops this handle
    c.close()


def rel_force_err(f, fref):
    """max_i |dF_i| / max_i |F_i| (the per-system force parity metric)."""
    d = np.linalg.norm(np.asarray(f) - np.asarray(fref), axis=1).max()
    s = np.linalg.norm(np.asarray(fref), axis=1).max()
    return d / s


def scaled_force_err(f, fref, scale):
    """max_i |dF_i| / scale_i, with scale_i the sum of the magnitudes of the
    force contributions summed into atom i (oracle.force_scale): the relative
    error of each atom's force sum.  Used along trajectories, where the net
    forces fall towards Fmax < 1e-3 eV/A while the contributions stay of order
    eV/A, so that max_i |F_i| stops measuring the arithmetic's precision."""
    d = np.linalg.norm(np.asarray(f) - np.asarray(fref), axis=1)
    return float((d / np.asarray(scale)).max())
