"""CPU checks of the C ABI library: it loads, exports every function that
include/*.h declares, its host-only logic is right, and the Python binding
fails loudly when the library is missing.  No device work here."""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lattice-dev_amd")
LIB = os.path.join(PKG, "liblattice_amd.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", PKG, "-j4"], check=True)
    return ctypes.CDLL(LIB)


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(lat_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_num_bins_matches_reference_formula(lib):
    lib.lat_num_bins.restype = ctypes.c_int
    for n in (8, 16, 64, 128, 192, 256, 1024):
        expect = int(np.floor(np.sqrt(3 * (n // 2 + 1) ** 2 + 1.0)))  # analysis_spec.f90:55-56
        assert lib.lat_num_bins(n, n, n) == expect
    assert lib.lat_num_bins(64, 64, 64) == 57 and lib.lat_num_bins(1024, 1024, 1024) == 888


def schedule(lib, time_var, order, dt, nsteps, tree=1):
    n = lib.lat_schedule(time_var, order, ctypes.c_double(dt), nsteps, tree, None, None, 0)
    assert n > 0
    t = (ctypes.c_int * n)()
    h = (ctypes.c_double * n)()
    assert lib.lat_schedule(time_var, order, ctypes.c_double(dt), nsteps, tree, t, h, n) == n
    return np.array(t[:]), np.array(h[:])


@pytest.mark.parametrize("time_var", [0, 1])
@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_schedule_invariants(lib, time_var, order):
    """Each of drift, grav and kick advances by exactly nsteps*dt in total; kicks
    are never adjacent; the sequence is a palindrome within each call."""
    dt, n = 0.01, 3
    t, h = schedule(lib, time_var, order, dt, n)
    for kind in (1, 2, 3):
        assert abs(h[t == kind].sum() - n * dt) < 1e-14
    k = np.nonzero(t == 3)[0]
    assert np.all(np.diff(k) > 1)
    stages = {2: 1, 4: 3, 6: 7, 8: 15}[order]
    assert len(k) == n * stages
    assert np.array_equal(t, t[::-1]) and np.allclose(h, h[::-1], rtol=0, atol=1e-15)
    # Hamiltonian_Split order: conformal 1 = drift (outermost), cosmic 1 = grav
    assert t[0] == (1 if time_var == 0 else 2)


def test_schedule_rejects_bad_order(lib):
    assert lib.lat_schedule(1, 5, ctypes.c_double(0.1), 1, 1, None, None, 0) < 0


def test_binding_fails_loudly_without_library(monkeypatch):
    import lattice_amd as la

    monkeypatch.setattr(la, "LIB_PATH", os.path.join(ROOT, "does-not-exist.so"))
    monkeypatch.setattr(la, "_lib", None)
    with pytest.raises(la.LatticeError):
        la.lib()
