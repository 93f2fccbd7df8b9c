"""ctypes front end of the CPU FP64 oracle (oracle/liblatoracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library
never calls it.

Array convention: a lattice field is a numpy array of shape (nz, ny, nx) in C
order, which is byte-identical to the reference's Fortran f(nx, ny, nz).  A
half-complex transform has shape (nz, ny, nx//2+1), complex128 (FFTW r2c).
Field sets are (nfld, nz, ny, nx).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblatoracle.so")
_lib = None

CONFORMAL, COSMIC = 0, 1
LAP_SPECTRAL, LAP_STENCIL = 0, 1
MODEL_PHI4, MODEL_TWOFIELD = 0, 1
E_NAMES = ["a", "hubble", "rho", "ke", "pe", "ge", "grav", "resid", "pmom",
           "mean1", "mean2", "ge_grad"]

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_stencil_grad2_sum.restype = ctypes.c_double
        _lib.orc_potential.restype = ctypes.c_double
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class Sim(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("nfld", ctypes.c_int),
        ("len", ctypes.c_double), ("time_var", ctypes.c_int), ("lap_kind", ctypes.c_int),
        ("c", ctypes.c_double * 5), ("model", ctypes.c_int), ("mp", ctypes.c_double * 4),
        ("fld", _dp), ("fldp", _dp), ("scratch", _dp),
        ("yscl", ctypes.c_double), ("ysclp", ctypes.c_double),
        ("gw", ctypes.c_int), ("gw_coup", ctypes.c_double), ("gwu", _dp), ("gwp", _dp),
    ]


# ---------------------------------------------------------------- FFT / operators

def fft_c2c_1d(x, sign=-1):
    x = _c128(x).copy()
    buf = x.view(np.float64)
    lib().orc_fft_c2c_1d(len(x), _p(buf), sign)
    return x


def r2c_3d(f):
    f = _f64(f)
    nz, ny, nx = f.shape
    out = np.zeros((nz, ny, nx // 2 + 1), np.complex128)
    lib().orc_r2c_3d(nx, ny, nz, _p(f), _p(out.view(np.float64)))
    return out


def c2r_3d(F, nx):
    F = _c128(F)
    nz, ny, nxh = F.shape
    out = np.zeros((nz, ny, nx))
    lib().orc_c2r_3d(nx, ny, nz, _p(F.view(np.float64)), _p(out))
    return out


def laplacian_3d(f, dk):
    f = _f64(f).copy()
    nz, ny, nx = f.shape
    lib().orc_laplacian_3d(nx, ny, nz, _p(f), ctypes.c_double(dk))
    return f


def gradient_squared_3d(f, dk):
    f = _f64(f)
    nz, ny, nx = f.shape
    out = np.zeros_like(f)
    lib().orc_gradient_squared_3d(nx, ny, nz, _p(f), _p(out), ctypes.c_double(dk))
    return out


def derivative_3d(f, dk, axis):
    """axis 0 = x, 1 = y, 2 = z (lattice axes, not numpy axes)."""
    f = _f64(f)
    nz, ny, nx = f.shape
    out = np.zeros_like(f)
    lib().orc_derivative_3d(nx, ny, nz, _p(f), _p(out), ctypes.c_double(dk), axis)
    return out


def stencil_laplacian(f, c, dx):
    f = _f64(f)
    nz, ny, nx = f.shape
    out = np.zeros_like(f)
    cc = (ctypes.c_double * 5)(*c)
    lib().orc_stencil_laplacian(nx, ny, nz, _p(f), _p(out), cc, ctypes.c_double(dx))
    return out


def stencil_grad2_sum(f, c):
    f = _f64(f)
    nz, ny, nx = f.shape
    cc = (ctypes.c_double * 5)(*c)
    return lib().orc_stencil_grad2_sum(nx, ny, nz, _p(f), cc)


def wrap_fields(g):
    """g: ghosted AoS array of shape (nz+2, ny+2, nx+2, nfld); filled in place."""
    assert g.flags["C_CONTIGUOUS"] and g.dtype == np.float64
    nzg, nyg, nxg, nfld = g.shape
    lib().orc_wrap_fields(nfld, nxg - 2, nyg - 2, nzg - 2, _p(g))
    return g


# ---------------------------------------------------------------- spectra

def num_bins(nx, ny, nz):
    return lib().orc_num_bins(nx, ny, nz)


def spectrum_3d(Fk, nx):
    Fk = _c128(Fk)
    nz, ny, _ = Fk.shape
    S = np.zeros(num_bins(nx, ny, nz))
    lib().orc_spectrum_3d(nx, ny, nz, _p(Fk.view(np.float64)), _p(S))
    return S


def crossspec_3d(Fk1, Fk2, nx):
    Fk1, Fk2 = _c128(Fk1), _c128(Fk2)
    nz, ny, _ = Fk1.shape
    ns = num_bins(nx, ny, nz)
    sre, sim = np.zeros(ns), np.zeros(ns)
    lib().orc_crossspec_3d(nx, ny, nz, _p(Fk1.view(np.float64)), _p(Fk2.view(np.float64)),
                           _p(sre), _p(sim))
    return sre, sim


def spectrum_discrete(Fk, n, kind):
    Fk = _c128(Fk)
    ns = int(1.7320508075688772935274463415059 * (n // 2) + 2.0)
    S = np.zeros(ns)
    lib().orc_spectrum_discrete(n, _p(Fk.view(np.float64)), _p(S), kind)
    return S


def spectrum_discrete_k(Fk, n, kind):
    Fk = _c128(Fk)
    ns = int(1.7320508075688772935274463415059 * (n // 2) + 2.0)
    S = np.zeros(ns)
    lib().orc_spectrum_discrete_k(n, _p(Fk.view(np.float64)), _p(S), kind)
    return S


def crossspec_discrete_k(Fk1, Fk2, n):
    Fk1, Fk2 = _c128(Fk1), _c128(Fk2)
    ns = int(1.7320508075688772935274463415059 * (n // 2) + 2.0)
    S1, S2 = np.zeros(ns), np.zeros(ns)
    lib().orc_crossspec_discrete_k(n, _p(Fk1.view(np.float64)), _p(Fk2.view(np.float64)), _p(S1), _p(S2))
    return S1, S2


# ---------------------------------------------------------------- evolution

class Simulation:
    """FP64 CPU evolution with the reference's split/composition semantics."""

    def __init__(self, n, nfld, length, time_var, lap_kind=LAP_SPECTRAL,
                 stencil=(-6.0, 1.0, 0.0, 0.0, 1.0), model=MODEL_TWOFIELD, mp=(1.0, 0.0), gw=False, gw_mpl=1.0):
        if isinstance(n, int):
            n = (n, n, n)
        self.nx, self.ny, self.nz = n
        self.nfld = nfld
        self.fld = np.zeros((nfld, self.nz, self.ny, self.nx))
        self.fldp = np.zeros_like(self.fld)
        self.scratch = np.zeros((self.nz, self.ny, self.nx))
        self.s = Sim()
        s = self.s
        s.nx, s.ny, s.nz, s.nfld = self.nx, self.ny, self.nz, nfld
        s.len = length
        s.time_var, s.lap_kind, s.model = time_var, lap_kind, model
        for i, v in enumerate(stencil):
            s.c[i] = v
        for i, v in enumerate(mp):
            s.mp[i] = v
        s.fld, s.fldp, s.scratch = _p(self.fld), _p(self.fldp), _p(self.scratch)
        s.yscl, s.ysclp = 1.0, 0.0
        # metric perturbations u_ij, components xx, xy, xz, yy, yz, zz
        self.gwu = np.zeros((6, self.nz, self.ny, self.nx))
        self.gwp = np.zeros_like(self.gwu)
        s.gw, s.gw_coup = int(bool(gw)), 2.0 / gw_mpl ** 2
        s.gwu, s.gwp = _p(self.gwu), _p(self.gwp)

    @property
    def yscl(self):
        return self.s.yscl

    @yscl.setter
    def yscl(self, v):
        self.s.yscl = v

    @property
    def ysclp(self):
        return self.s.ysclp

    @ysclp.setter
    def ysclp(self, v):
        self.s.ysclp = v

    def set_state(self, fld, fldp, yscl, ysclp):
        self.fld[...] = fld
        self.fldp[...] = fldp
        self.s.yscl, self.s.ysclp = yscl, ysclp

    def split(self, h, ind):
        rc = lib().orc_split(ctypes.byref(self.s), ctypes.c_double(h), ind)
        assert rc == 0

    def symp_o2step(self, dt, w1, w2):
        lib().orc_symp_o2step(ctypes.byref(self.s), ctypes.c_double(dt), ctypes.c_double(w1),
                              ctypes.c_double(w2))

    def symp(self, order, dt, nsteps, symp6_tree=1):
        rc = lib().orc_symp(ctypes.byref(self.s), order, ctypes.c_double(dt), nsteps, symp6_tree)
        assert rc == 0

    def calc_metric(self):
        lib().orc_calc_metric(ctypes.byref(self.s))

    def energies(self):
        out = np.zeros(len(E_NAMES))
        lib().orc_energies(ctypes.byref(self.s), _p(out))
        return dict(zip(E_NAMES, out))

    def discrete_output(self, time):
        """Discrete-tree make_output quantities: (LOG.out line values,
        lnrho field, dlnrho field, PSD columns (4, ns_d))."""
        n = self.nx
        ns = int(1.7320508075688772935274463415059 * (n // 2) + 2.0)
        log = np.zeros(12 + 2 * self.nfld)
        lnr, dln = np.zeros((n, n, n)), np.zeros((n, n, n))
        psd = np.zeros((4, ns))
        rc = lib().orc_discrete_output(ctypes.byref(self.s), ctypes.c_double(time), _p(log), _p(lnr), _p(dln),
                                       _p(psd))
        assert rc > 0
        return log, lnr, dln, psd


# ---------------------------------------------------------------- GW

def gw_source_fd(fld, dx):
    fld = _f64(fld)
    nfld, nz, ny, nx = fld.shape
    out = np.zeros((6, nz, ny, nx))
    lib().orc_gw_source_fd(nx, ny, nz, nfld, _p(fld), ctypes.c_double(dx), _p(out))
    return out


def gw_source(fld, dk):
    fld = _f64(fld)
    nfld, nz, ny, nx = fld.shape
    out = np.zeros((6, nz, ny, nx))
    lib().orc_gw_source(nx, ny, nz, nfld, _p(fld), ctypes.c_double(dk), _p(out))
    return out


def tt_project(Uk, nx):
    Uk = _c128(Uk).copy()
    _, nz, ny, _ = Uk.shape
    lib().orc_tt_project(nx, ny, nz, _p(Uk.view(np.float64)))
    return Uk


def gw_spectrum(dUk, nx, dk=1.0):
    dUk = _c128(dUk)
    _, nz, ny, _ = dUk.shape
    S = np.zeros(num_bins(nx, ny, nz))
    lib().orc_gw_spectrum(nx, ny, nz, None, _p(dUk.view(np.float64)), ctypes.c_double(dk), _p(S))
    return S


# ---------------------------------------------------------------- IC

def philox_uniforms(seed, stream, count):
    u1, u2 = np.zeros(count), np.zeros(count)
    lib().orc_philox_normal_pairs(ctypes.c_ulonglong(seed), ctypes.c_ulonglong(stream),
                                  ctypes.c_ulonglong(count), _p(u1), _p(u2))
    return u1, u2


def sample(n, length, gamma, m2eff, mpl, seed, stream, kcut_fac=1.0):
    if isinstance(n, int):
        n = (n, n, n)
    nx, ny, nz = n
    out = np.zeros((nz, ny, nx))
    lib().orc_sample(nx, ny, nz, ctypes.c_double(length), ctypes.c_double(gamma),
                     ctypes.c_double(m2eff), ctypes.c_double(mpl), ctypes.c_double(kcut_fac),
                     ctypes.c_ulonglong(seed), ctypes.c_ulonglong(stream), _p(out))
    return out


def dst2(x):
    """FFTW RODFT10 (DST-II), unnormalised: Y_k = 2 sum_j X_j sin(pi (j+1/2)(k+1)/n)."""
    x = _f64(x).copy()
    lib().orc_dst2(len(x), _p(x))
    return x


def sample_kernel(n, length, gamma, m2eff, mpl, kcut_fac=1.0):
    """The sampler's real-space convolution kernel (before the noise)."""
    if isinstance(n, int):
        n = (n, n, n)
    nx, ny, nz = n
    out = np.zeros((nz, ny, nx))
    lib().orc_sample_kernel(nx, ny, nz, ctypes.c_double(length), ctypes.c_double(gamma), ctypes.c_double(m2eff),
                            ctypes.c_double(mpl), ctypes.c_double(kcut_fac), _p(out))
    return out
