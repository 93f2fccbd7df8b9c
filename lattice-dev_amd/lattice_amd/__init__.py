"""Host-side binding of liblattice_amd.so (the C ABI of include/lattice_c.h).

This mirrors the reference's Fortran module interfaces for the evolution path
(`module fftw3`, `module analysis`, `module Hamiltonian`, `module integrator`):
same names, argument meaning and error behaviour (a nonzero status raises
LatticeError, where the reference prints and stops).  It holds no numerics of
its own: every call goes through the HIP library, and importing fails loudly
if that library is missing.

Array convention: a lattice field is a numpy array of shape (nz, ny, nx),
C-contiguous, i.e. byte-identical to the reference's Fortran f(nx, ny, nz).
Half-complex transforms are (nz, ny, nx//2+1) complex128 (FFTW r2c layout).
Field sets are (nfld, nz, ny, nx) here and are converted to/from the
reference's AoS fld(nfld, nx, ny, nz) layout at the C boundary.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "liblattice_amd.so")

F64, MIXED = 0, 1
LAP_SPECTRAL, LAP_STENCIL = 0, 1
CONFORMAL, COSMIC = 0, 1
MODEL_PHI4, MODEL_TWOFIELD = 0, 1
E_NAMES = ["a", "hubble", "rho", "ke", "pe", "ge", "grav", "resid", "pmom",
           "mean1", "mean2", "ge_grad"]

STENCIL_7PT = (-6.0, 1.0, 0.0, 0.0, 1.0)      # Discrete/Entropy/params.f90:63
STENCIL_27PT = (-128.0, 14.0, 3.0, 1.0, 30.0)  # hamiltonian_conformal.f90:41

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


class LatticeError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("nfld", ctypes.c_int),
        ("len", ctypes.c_double), ("precision", ctypes.c_int), ("laplacian", ctypes.c_int),
        ("stencil", ctypes.c_double * 5), ("time_var", ctypes.c_int), ("model", ctypes.c_int),
        ("model_params", ctypes.c_double * 4), ("symp6_tree", ctypes.c_int),
        ("gw", ctypes.c_int), ("gw_mpl", ctypes.c_double), ("reserved", ctypes.c_int * 4),
    ]


EXPORTS = [
    "lat_create", "lat_destroy", "lat_last_error", "lat_synchronize",
    "lat_laplacian_3d", "lat_gradient_squared_3d", "lat_fft_r2c_3d", "lat_fft_c2r_3d",
    "lat_stencil_laplacian", "lat_num_bins", "lat_spectrum_3d", "lat_crossspec_3d",
    "lat_output_spectra", "lat_set_state", "lat_get_state", "lat_get_scalars",
    "lat_set_scalars", "lat_split", "lat_symp_o2step", "lat_symp", "lat_calc_metric",
    "lat_energies", "lat_sample_field", "lat_device_state", "lat_profile", "lat_profile_read",
    "lat_num_bins_discrete", "lat_spectrum_discrete",
    "lat_set_gw", "lat_get_gw", "lat_gw_source", "lat_gw_spectrum",
    "lat_spectrum_discrete_k", "lat_crossspec_discrete_k",
    "lat_create_batch", "lat_num_members", "lat_member",
    "lat_ensemble_create", "lat_ensemble_destroy", "lat_ensemble_last_error", "lat_ensemble_size",
    "lat_ensemble_member", "lat_ensemble_use_graph", "lat_ensemble_single_launch", "lat_ensemble_symp",
    "lat_ensemble_synchronize", "lat_ensemble_set_tuning", "lat_ensemble_tuning",
    "lat_autotune", "lat_tuning", "lat_set_tuning", "lat_copy_bandwidth", "lat_make_output",
    "lat_set_laplacian_form", "lat_laplacian_form", "lat_set_kspace_order", "lat_place_tune",
    "lat_checkpoint", "lat_restore",
    "lat_corrector_create", "lat_corrector_destroy", "lat_corrector_last_error", "lat_corrector_set_layer",
    "lat_corrector_get_layer", "lat_corrector_set_norm", "lat_corrector_get_norm", "lat_corrector_infer",
    "lat_corrector_gradients", "lat_corrector_train", "lat_corrector_step", "lat_corrector_step_judged",
]
STAGES = ["x_r2c", "y_fwd", "z_lap", "y_inv", "x_c2r", "scalar", "drift", "stencil", "reduce", "gw_src",
          "sep_x", "sep_y", "sep_z"]


def lib():
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LatticeError(f"{LIB_PATH} not built: run __graft_entry__.build() "
                               "(make -C lattice-dev_amd)")
        L = ctypes.CDLL(LIB_PATH)
        L.lat_last_error.restype = ctypes.c_char_p
        L.lat_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_void_p)]
        L.lat_destroy.argtypes = [ctypes.c_void_p]
        for name in EXPORTS:
            fn = getattr(L, name)
            if name not in ("lat_last_error", "lat_destroy", "lat_create", "lat_ensemble_destroy",
                            "lat_ensemble_last_error", "lat_corrector_destroy", "lat_corrector_last_error"):
                fn.restype = ctypes.c_int
        L.lat_corrector_last_error.restype = ctypes.c_char_p
        L.lat_corrector_destroy.argtypes = [ctypes.c_void_p]
        L.lat_ensemble_last_error.restype = ctypes.c_char_p
        L.lat_ensemble_destroy.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _p(a):
    if not a.flags["C_CONTIGUOUS"]:
        raise LatticeError("array must be C-contiguous")
    return a.ctypes.data_as(_dp)


def num_bins(nx, ny, nz):
    """init_spectrum_3d (analysis_spec.f90:48-57): number of shells ns."""
    return lib().lat_num_bins(nx, ny, nz)


class Lattice:
    """One device context: state, FFT tables, stream (the module globals of
    hamiltonian_*.f90 and fftw_mod.f90 gathered into an object)."""

    def __init__(self, n, nfld=2, length=10.0, precision=F64, laplacian=LAP_SPECTRAL,
                 stencil=STENCIL_7PT, time_var=COSMIC, model=MODEL_TWOFIELD, model_params=(1.0, 2.5e5),
                 symp6_tree=1, gw=False, gw_mpl=1.0, members=None):
        if isinstance(n, int):
            n = (n, n, n)
        self.nx, self.ny, self.nz = n
        self.nfld = nfld
        self.length = length
        self.dx = length / self.nx
        self.dk = 2.0 * np.pi / length
        cfg = Config()
        cfg.nx, cfg.ny, cfg.nz, cfg.nfld = self.nx, self.ny, self.nz, nfld
        cfg.len, cfg.precision, cfg.laplacian = length, precision, laplacian
        for i, v in enumerate(stencil):
            cfg.stencil[i] = v
        cfg.time_var, cfg.model = time_var, model
        for i, v in enumerate(model_params):
            cfg.model_params[i] = v
        cfg.symp6_tree = symp6_tree
        cfg.gw, cfg.gw_mpl = int(bool(gw)), gw_mpl
        self.cfg = cfg
        h = ctypes.c_void_p()
        if members is None:
            rc = lib().lat_create(ctypes.byref(cfg), ctypes.byref(h))
        else:
            # a batch: `members` = list of per-member model_params tuples
            mp = np.zeros((len(members), 4))
            for m, par in enumerate(members):
                mp[m, :len(par)] = par
            self._mp = mp
            rc = lib().lat_create_batch(ctypes.byref(cfg), len(members), _p(mp), ctypes.byref(h))
        if rc != 0:
            raise LatticeError(f"lat_create failed with status {rc}")
        self._h = h
        self._parent = None
        self.ns = num_bins(self.nx, self.ny, self.nz)

    def autotune(self):
        """Pick the faster launch shape of each sweep (sizes up to 256)."""
        self._check(lib().lat_autotune(self._h), "lat_autotune")
        return self.tuning()

    def place_tune(self, candidates=4):
        """Keep the fastest of `candidates` placements of the field, momentum
        and k-space arrays in HBM, the state unchanged (lat_place_tune).
        Returns (index of the set kept, ms per force evaluation of each set;
        set 0 = the arrays the context had, -1 = not allocated)."""
        ms = (ctypes.c_float * max(int(candidates), 1))()
        r = lib().lat_place_tune(self._h, int(candidates), ms)
        if r < 0:
            self._check(r, "lat_place_tune")
        return r, [float(x) for x in ms]

    def copy_bandwidth(self, nbytes=2 << 30, reps=5):
        """Measured device-copy bandwidth in GB/s (read + write)."""
        out = ctypes.c_double()
        self._check(lib().lat_copy_bandwidth(self._h, ctypes.c_size_t(nbytes), reps, ctypes.byref(out)),
                    "lat_copy_bandwidth")
        return out.value

    def set_laplacian_form(self, separable):
        """False: five-sweep 3-D spectral Laplacian; True: the same operator
        as a sum of one-axis second derivatives in three sweeps (mixed form)."""
        self._check(lib().lat_set_laplacian_form(self._h, int(bool(separable))), "lat_set_laplacian_form")

    def laplacian_form(self):
        return lib().lat_laplacian_form(self._h)

    def set_kspace_order(self, blocked):
        """Row order of the force pipeline's k-space arrays: True = blocked
        (cubic 256 and 1024), False = FFTW order (default)."""
        self._check(lib().lat_set_kspace_order(self._h, int(bool(blocked))), "lat_set_kspace_order")

    def set_tuning(self, x_r2c=0, y=0, z=0, x_c2r=0):
        vals = (ctypes.c_int * 4)(x_r2c, y, z, x_c2r)
        self._check(lib().lat_set_tuning(self._h, vals), "lat_set_tuning")

    def tuning(self):
        out = (ctypes.c_int * 4)()
        self._check(lib().lat_tuning(self._h, out), "lat_tuning")
        return dict(zip(("x_r2c", "y", "z", "x_c2r"), list(out)))

    @property
    def members(self):
        return lib().lat_num_members(self._h)

    def member(self, m):
        """View of member m of a batch: the per-member API (state, IC,
        energies, spectra) on that realisation.  Keeps the batch alive."""
        h = ctypes.c_void_p()
        self._check(lib().lat_member(self._h, m, ctypes.byref(h)), "lat_member")
        v = object.__new__(Lattice)
        v.__dict__.update({k: val for k, val in self.__dict__.items() if k not in ("_h", "_parent", "_mp")})
        v._h = h
        v._parent = self
        return v

    # ---------------------------------------------------------------- plumbing
    def close(self):
        if getattr(self, "_h", None):
            lib().lat_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().lat_last_error(self._h)
            raise LatticeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def _field(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        if f.shape != (self.nz, self.ny, self.nx):
            raise LatticeError(f"field shape {f.shape} != {(self.nz, self.ny, self.nx)}")
        return f

    def synchronize(self):
        self._check(lib().lat_synchronize(self._h), "lat_synchronize")

    # ---------------------------------------------------------------- module fftw3
    def laplacian_3d(self, f, dk=None):
        """laplacian_3d (fftw_mod.f90:177-207); returns a new array."""
        f = self._field(f).copy()
        dk = self.dk if dk is None else dk
        self._check(lib().lat_laplacian_3d(self._h, self.nx, self.ny, self.nz, _p(f), ctypes.c_double(dk)),
                    "laplacian_3d")
        return f

    def gradient_squared_3d(self, f, dk=None):
        f = self._field(f)
        out = np.empty_like(f)
        dk = self.dk if dk is None else dk
        self._check(lib().lat_gradient_squared_3d(self._h, _p(f), _p(out), ctypes.c_double(dk)),
                    "gradient_squared_3d")
        return out

    def fft_r2c_3d(self, f):
        f = self._field(f)
        F = np.empty((self.nz, self.ny, self.nx // 2 + 1), np.complex128)
        self._check(lib().lat_fft_r2c_3d(self._h, _p(f), _p(F.view(np.float64))), "fft_r2c_3d")
        return F

    def fft_c2r_3d(self, F):
        F = np.ascontiguousarray(F, dtype=np.complex128)
        f = np.empty((self.nz, self.ny, self.nx))
        self._check(lib().lat_fft_c2r_3d(self._h, _p(F.view(np.float64)), _p(f)), "fft_c2r_3d")
        return f

    def stencil_laplacian(self, f):
        f = self._field(f)
        out = np.empty_like(f)
        self._check(lib().lat_stencil_laplacian(self._h, _p(f), _p(out)), "stencil_laplacian")
        return out

    # ---------------------------------------------------------------- module analysis
    def spectrum_3d(self, f, return_fk=False):
        """spectrum_3d (analysis_spec.f90:59-102)."""
        f = self._field(f)
        S = np.empty(self.ns)
        F = np.empty((self.nz, self.ny, self.nx // 2 + 1), np.complex128) if return_fk else None
        self._check(lib().lat_spectrum_3d(self._h, _p(S), _p(f), _p(F.view(np.float64)) if return_fk else None),
                    "spectrum_3d")
        return (S, F) if return_fk else S

    def spectrum_discrete(self, f, kind=0):
        """Discrete-tree spectra (cubic lattice): kind 0 = spectrum(f, S)
        (Discrete/Entropy/analysis.f90:89-114, two-bin weights, /n^6);
        kind 1 = spectrum_bin(f, S) (analysis.f90:22-52, integer counts)."""
        f = self._field(f)
        S = np.empty(lib().lat_num_bins_discrete(self.nx))
        self._check(lib().lat_spectrum_discrete(self._h, _p(f), _p(S), int(kind)), "spectrum_discrete")
        return S

    def spectrum_discrete_k(self, Fk, kind=0):
        """kind 0 = spectrum_nofft (Discrete/Entropy/analysis.f90:118-147),
        kind 1 = spectrum_bin_nofft (:54-84), on a (n, n, n/2+1) transform."""
        Fk = np.ascontiguousarray(Fk, dtype=np.complex128)
        S = np.empty(lib().lat_num_bins_discrete(self.nx))
        self._check(lib().lat_spectrum_discrete_k(self._h, _p(Fk.view(np.float64)), _p(S), int(kind)),
                    "spectrum_discrete_k")
        return S

    def spectrum_discrete_k_noalloc(self, fk_dev_ptr, s_dev_ptr, kind=0):
        """lat_spectrum_discrete_k on a device-resident transform into device
        memory, allocating nothing (pointers are device addresses)."""
        self._check(lib().lat_spectrum_discrete_k_noalloc(self._h, ctypes.c_void_p(fk_dev_ptr),
                                                          ctypes.c_void_p(s_dev_ptr), kind),
                    "spectrum_discrete_k_noalloc")

    def crossspec_discrete_k(self, Fk1, Fk2):
        """crossspec_nofft (Discrete/Entropy/analysis.f90:149-186) -> (S1, S2)."""
        Fk1 = np.ascontiguousarray(Fk1, dtype=np.complex128)
        Fk2 = np.ascontiguousarray(Fk2, dtype=np.complex128)
        ns = lib().lat_num_bins_discrete(self.nx)
        s1, s2 = np.empty(ns), np.empty(ns)
        self._check(lib().lat_crossspec_discrete_k(self._h, _p(Fk1.view(np.float64)), _p(Fk2.view(np.float64)),
                                                   _p(s1), _p(s2)), "crossspec_discrete_k")
        return s1, s2

    def crossspec_3d(self, Fk1, Fk2):
        """crossspec_3d (analysis_spec.f90:145-178) -> (S_real, S_imag)."""
        Fk1 = np.ascontiguousarray(Fk1, dtype=np.complex128)
        Fk2 = np.ascontiguousarray(Fk2, dtype=np.complex128)
        sre, sim = np.empty(self.ns), np.empty(self.ns)
        self._check(lib().lat_crossspec_3d(self._h, _p(Fk1.view(np.float64)), _p(Fk2.view(np.float64)),
                                           _p(sre), _p(sim)), "crossspec_3d")
        return sre, sim

    def output_spectra(self, field=0):
        """make_output spectra (evolve_spectral.f90:531-538) of the resident state:
        returns (S_phi, S_pi, Re S_pi_phi, Im S_pi_phi)."""
        out = np.empty(4 * self.ns)
        self._check(lib().lat_output_spectra(self._h, field, _p(out)), "output_spectra")
        return out.reshape(4, self.ns)

    # ---------------------------------------------------------------- gravitational waves
    # u_ij as (6, nz, ny, nx) arrays, components xx, xy, xz, yy, yz, zz; the C ABI
    # takes the reference's hij(2,3,nx,ny,nz) (grav-waves/data-gw.f90:28).
    def _gw_to_aos(self, u):
        u = np.asarray(u, np.float64)
        aos = np.empty((self.nz, self.ny, self.nx, 3, 2))
        aos[..., 0, 0], aos[..., 1, 0], aos[..., 2, 0] = u[0], u[3], u[5]
        aos[..., 0, 1], aos[..., 1, 1], aos[..., 2, 1] = u[1], u[2], u[4]
        return aos

    def _gw_from_aos(self, aos):
        return np.stack([aos[..., 0, 0], aos[..., 0, 1], aos[..., 1, 1], aos[..., 1, 0], aos[..., 2, 1],
                         aos[..., 2, 0]])

    def set_gw(self, u, pu):
        """Upload u_ij and their momenta, (6, nz, ny, nx) each."""
        a, b = self._gw_to_aos(u), self._gw_to_aos(pu)
        self._check(lib().lat_set_gw(self._h, _p(a), _p(b)), "set_gw")

    def set_gw_aos(self, hij, pi_hij):
        """Reference layout hij(2,3,nx,ny,nz) as a C-order (nz, ny, nx, 3, 2) array."""
        a = np.ascontiguousarray(hij, np.float64)
        b = np.ascontiguousarray(pi_hij, np.float64)
        self._check(lib().lat_set_gw(self._h, _p(a), _p(b)), "set_gw")

    def get_gw(self):
        a = np.empty((self.nz, self.ny, self.nx, 3, 2))
        b = np.empty_like(a)
        self._check(lib().lat_get_gw(self._h, _p(a), _p(b)), "get_gw")
        return self._gw_from_aos(a), self._gw_from_aos(b)

    def gw_source(self):
        """S_ij = sum_l d_i phi_l d_j phi_l (central differences) of the resident fields."""
        out = np.empty((6, self.nz, self.ny, self.nx))
        self._check(lib().lat_gw_source(self._h, _p(out)), "gw_source")
        return out

    def gw_spectrum(self):
        """(sum_ij |Lambda pi_u^|^2, sum_ij |Lambda u^|^2) binned per shell."""
        out = np.empty(2 * self.ns)
        self._check(lib().lat_gw_spectrum(self._h, _p(out)), "gw_spectrum")
        return out.reshape(2, self.ns)

    def discrete_output(self, time, dt=0.0, mpl=1.0, seed=0, directory=None):
        """Discrete-tree make_output (iofile.f90:317-335): returns the LOG.out
        values and the PSD columns (4, ns_d); with `directory`, appends LOG.out
        and PSD there."""
        vals = np.empty(12 + 2 * self.nfld)
        ns = lib().lat_num_bins_discrete(self.nx)
        psd = np.empty((4, ns))
        d = directory.encode() if directory is not None else None
        self._check(lib().lat_discrete_output(self._h, ctypes.c_double(time), ctypes.c_double(dt),
                                              ctypes.c_double(mpl), int(seed), d, _p(vals), _p(psd)),
                    "discrete_output")
        return vals, psd

    def make_output(self, time, directory=".", tree=0):
        """make_output/dump_rho: append to energy_spec.out and spectrum.out."""
        self._check(lib().lat_make_output(self._h, ctypes.c_double(time), int(tree),
                                          os.fsencode(str(directory))), "make_output")

    def checkpoint(self, directory, time, outframenum):
        self._check(lib().lat_checkpoint(self._h, os.fsencode(str(directory)), ctypes.c_double(time),
                                         int(outframenum)), "checkpoint")

    def restore(self, directory):
        """Returns (time, outframenum) of the checkpoint."""
        t, fr = ctypes.c_double(), ctypes.c_int()
        self._check(lib().lat_restore(self._h, os.fsencode(str(directory)), ctypes.byref(t), ctypes.byref(fr)),
                    "restore")
        return t.value, fr.value

    # ---------------------------------------------------------------- state
    def set_state(self, fld, fldp, yscl=1.0, ysclp=0.0):
        """fld, fldp: (nfld, nz, ny, nx) SoA arrays; converted to the reference AoS."""
        fld = np.asarray(fld, np.float64)
        fldp = np.asarray(fldp, np.float64)
        aos = np.ascontiguousarray(np.moveaxis(fld, 0, -1))
        aosp = np.ascontiguousarray(np.moveaxis(fldp, 0, -1))
        self._check(lib().lat_set_state(self._h, _p(aos), _p(aosp), 0, ctypes.c_double(yscl),
                                        ctypes.c_double(ysclp)), "set_state")

    def set_state_aos(self, fld_aos, fldp_aos, ghost, yscl, ysclp):
        """Reference layout: fld(nfld, 0:nx+1, ...) (ghost=1) and fldp(nfld, 1:nx, ...)."""
        fld_aos = np.ascontiguousarray(fld_aos, np.float64)
        fldp_aos = np.ascontiguousarray(fldp_aos, np.float64)
        self._check(lib().lat_set_state(self._h, _p(fld_aos), _p(fldp_aos), ghost, ctypes.c_double(yscl),
                                        ctypes.c_double(ysclp)), "set_state")

    def get_state(self):
        aos = np.empty((self.nz, self.ny, self.nx, self.nfld))
        aosp = np.empty_like(aos)
        y, p = ctypes.c_double(), ctypes.c_double()
        self._check(lib().lat_get_state(self._h, _p(aos), _p(aosp), 0, ctypes.byref(y), ctypes.byref(p)),
                    "get_state")
        return np.moveaxis(aos, -1, 0).copy(), np.moveaxis(aosp, -1, 0).copy(), y.value, p.value

    def get_state_aos(self, ghost):
        g = ghost
        aos = np.empty((self.nz + 2 * g, self.ny + 2 * g, self.nx + 2 * g, self.nfld))
        aosp = np.empty((self.nz, self.ny, self.nx, self.nfld))
        y, p = ctypes.c_double(), ctypes.c_double()
        self._check(lib().lat_get_state(self._h, _p(aos), _p(aosp), g, ctypes.byref(y), ctypes.byref(p)),
                    "get_state")
        return aos, aosp, y.value, p.value

    @property
    def scalars(self):
        y, p = ctypes.c_double(), ctypes.c_double()
        self._check(lib().lat_get_scalars(self._h, ctypes.byref(y), ctypes.byref(p)), "get_scalars")
        return y.value, p.value

    def set_scalars(self, yscl, ysclp):
        self._check(lib().lat_set_scalars(self._h, ctypes.c_double(yscl), ctypes.c_double(ysclp)), "set_scalars")

    # ---------------------------------------------------------------- module Hamiltonian / integrator
    def Hamiltonian_Split(self, dt, hamind):
        self._check(lib().lat_split(self._h, ctypes.c_double(dt), hamind), "Hamiltonian_Split")

    def symp_o2step(self, dt, w1, w2):
        self._check(lib().lat_symp_o2step(self._h, ctypes.c_double(dt), ctypes.c_double(w1), ctypes.c_double(w2)),
                    "symp_o2step")

    def symp(self, order, dt, nsteps):
        self._check(lib().lat_symp(self._h, order, ctypes.c_double(dt), nsteps), f"symp{order}")

    def symp2(self, dt, nsteps):
        self.symp(2, dt, nsteps)

    def symp4(self, dt, nsteps):
        self.symp(4, dt, nsteps)

    def symp6(self, dt, nsteps):
        self.symp(6, dt, nsteps)

    def symp8(self, dt, nsteps):
        self.symp(8, dt, nsteps)

    def calc_metric(self):
        self._check(lib().lat_calc_metric(self._h), "calc_metric")

    def energies(self):
        out = np.empty(len(E_NAMES))
        self._check(lib().lat_energies(self._h, _p(out)), "energies")
        return dict(zip(E_NAMES, out))

    def device_state(self):
        f, p = ctypes.c_void_p(), ctypes.c_void_p()
        self._check(lib().lat_device_state(self._h, ctypes.byref(f), ctypes.byref(p)), "device_state")
        return f.value, p.value

    # ---------------------------------------------------------------- initial conditions
    def sample_field(self, field, momentum, mean, gamma, m2eff, mpl, seed, stream, kcut_fac=1.0):
        """sample() + init_fields assignment (evolve_spectral.f90:262-303,421-506) on the device."""
        self._check(lib().lat_sample_field(self._h, field, momentum, ctypes.c_double(mean), ctypes.c_double(gamma),
                                           ctypes.c_double(m2eff), ctypes.c_double(mpl), ctypes.c_double(kcut_fac),
                                           ctypes.c_ulonglong(seed), ctypes.c_ulonglong(stream)), "sample_field")

    # ---------------------------------------------------------------- measurement
    def profile(self, enable=True):
        self._check(lib().lat_profile(self._h, 1 if enable else 0), "profile")

    def profile_read(self):
        ms = (ctypes.c_double * len(STAGES))()
        cnt = (ctypes.c_longlong * len(STAGES))()
        self._check(lib().lat_profile_read(self._h, ms, cnt), "profile_read")
        return {s: (ms[i], cnt[i]) for i, s in enumerate(STAGES)}


class Ensemble:
    """Realisations of uneven lattice sizes advanced together (one batched
    launch per sweep per size class, Yoshida calls replayed from a hipGraph).
    classes: list of (n, count, length); members: per-member model_params in
    class-major order (len = total count), or None for `model_params`."""

    def __init__(self, classes, nfld=2, precision=F64, laplacian=LAP_SPECTRAL, stencil=STENCIL_7PT,
                 time_var=COSMIC, model=MODEL_TWOFIELD, model_params=(1.0, 2.5e5), members=None):
        cfg = Config()
        cfg.nx = cfg.ny = cfg.nz = classes[0][0]
        cfg.nfld, cfg.len, cfg.precision, cfg.laplacian = nfld, classes[0][2], precision, laplacian
        for i, v in enumerate(stencil):
            cfg.stencil[i] = v
        cfg.time_var, cfg.model = time_var, model
        for i, v in enumerate(model_params):
            cfg.model_params[i] = v
        cfg.symp6_tree = 1
        self.classes = [tuple(c) for c in classes]
        n = (ctypes.c_int * len(classes))(*[c[0] for c in classes])
        cnt = (ctypes.c_int * len(classes))(*[c[1] for c in classes])
        ln = np.array([c[2] for c in classes], np.float64)
        mp = None
        if members is not None:
            mp = np.zeros((len(members), 4))
            for m, par in enumerate(members):
                mp[m, :len(par)] = par
        self._mp = mp
        h = ctypes.c_void_p()
        rc = lib().lat_ensemble_create(ctypes.byref(cfg), len(classes), n, cnt, _p(ln),
                                       _p(mp) if mp is not None else None, ctypes.byref(h))
        if rc != 0:
            raise LatticeError(f"lat_ensemble_create failed with status {rc}")
        self._h = h
        self._tmpl = dict(nfld=nfld)

    def close(self):
        if getattr(self, "_h", None):
            lib().lat_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().lat_ensemble_last_error(self._h)
            raise LatticeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    @property
    def size(self):
        return lib().lat_ensemble_size(self._h)

    def member(self, index):
        """Lattice view of member `index` (class-major)."""
        h = ctypes.c_void_p()
        self._check(lib().lat_ensemble_member(self._h, index, ctypes.byref(h)), "lat_ensemble_member")
        k = 0
        first = 0
        while index >= first + self.classes[k][1]:
            first += self.classes[k][1]
            k += 1
        nn, _, length = self.classes[k]
        v = object.__new__(Lattice)
        v.nx = v.ny = v.nz = nn
        v.nfld = self._tmpl["nfld"]
        v.length = length
        v.dx = length / nn
        v.dk = 2.0 * np.pi / length
        v.ns = num_bins(nn, nn, nn)
        v._h = h
        v._parent = self
        return v

    def use_graph(self, enable):
        self._check(lib().lat_ensemble_use_graph(self._h, int(bool(enable))), "use_graph")

    def set_tuning(self, k, shape):
        """Launch shapes (x R2C, y, z, x C2R; 0/1 each) of size class k."""
        arr = (ctypes.c_int * 4)(*[int(v) for v in shape])
        self._check(lib().lat_ensemble_set_tuning(self._h, k, arr), "set_tuning")

    def tuning(self, k):
        arr = (ctypes.c_int * 4)()
        self._check(lib().lat_ensemble_tuning(self._h, k, arr), "tuning")
        return tuple(arr)

    def single_launch(self, enable=None):
        """Force one launch per FFT sweep for all size classes (True) or one
        per class (False); None only queries the mode the autotuner picked.
        Returns whether the single launch is in use."""
        rc = lib().lat_ensemble_single_launch(self._h, -1 if enable is None else int(bool(enable)))
        if rc < 0:
            self._check(rc, "single_launch")
        return rc == 1

    def symp(self, order, dt, nsteps):
        """dt: one time step per class."""
        dt = np.ascontiguousarray(dt, np.float64)
        self._check(lib().lat_ensemble_symp(self._h, order, _p(dt), nsteps), "lat_ensemble_symp")

    def synchronize(self):
        self._check(lib().lat_ensemble_synchronize(self._h), "lat_ensemble_synchronize")


_fp = ctypes.POINTER(ctypes.c_float)


def _pf(a):
    return a.ctypes.data_as(_fp)


class Corrector:
    """The learned corrector (lat_corrector_*): a 4-layer BF16 3-D conv net
    on MFMA trained on the GPU on pairs from the plain integrator."""

    SHAPES = [(8, 32), (32, 32), (32, 32), (32, 4)]   # (cin, cout) per layer
    # input channels: the fluctuations (phi, chi, pi_phi, pi_chi) over their
    # spatial standard deviations, then the uniform <phi>, <pi_phi>,
    # log10(g^2/m^2) - 5 and log10(s_chi) / 10

    def __init__(self, seed=1):
        h = ctypes.c_void_p()
        rc = lib().lat_corrector_create(ctypes.c_ulonglong(seed), ctypes.byref(h))
        if rc != 0:
            raise LatticeError(f"lat_corrector_create failed with status {rc}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().lat_corrector_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().lat_corrector_last_error(self._h)
            raise LatticeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_layer(self, layer, w, b):
        """w: (cout, 27, cin) with tap = (dz+1)*9 + (dy+1)*3 + (dx+1); b: (cout,)."""
        w = np.ascontiguousarray(w, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        self._check(lib().lat_corrector_set_layer(self._h, layer, _pf(w), _pf(b)), "set_layer")

    def get_layer(self, layer):
        cin, cout = self.SHAPES[layer]
        w = np.empty((cout, 27, cin), np.float32)
        b = np.empty(cout, np.float32)
        self._check(lib().lat_corrector_get_layer(self._h, layer, _pf(w), _pf(b)), "get_layer")
        return w, b

    def set_norm(self, ratio):
        """Per-channel residual scale: RMS of the residual fluctuation over the
        standard deviation of the coarse state (phi, chi, pi_phi, pi_chi)."""
        r = np.ascontiguousarray(ratio, np.float64)
        self._check(lib().lat_corrector_set_norm(self._h, _p(r)), "set_norm")

    def get_norm(self):
        r = np.empty(4)
        self._check(lib().lat_corrector_get_norm(self._h, _p(r)), "get_norm")
        return r

    def infer(self, lat):
        out = np.empty((lat.nz, lat.ny, lat.nx, 4), np.float32)
        self._check(lib().lat_corrector_infer(self._h, lat._h, _pf(out)), "infer")
        return out

    def gradients(self, lat, target):
        """target: (nz, ny, nx, 4) float32; returns (loss, [(dW, db) per layer])."""
        target = np.ascontiguousarray(target, np.float32)
        total = sum(cout * 27 * cin + cout for cin, cout in self.SHAPES)
        g = np.empty(total, np.float32)
        loss = ctypes.c_double()
        self._check(lib().lat_corrector_gradients(self._h, lat._h, _pf(target), ctypes.byref(loss), _pf(g)),
                    "gradients")
        out, off = [], 0
        for cin, cout in self.SHAPES:
            nw = cout * 27 * cin
            out.append((g[off:off + nw].reshape(cout, 27, cin), g[off + nw:off + nw + cout]))
            off += nw + cout
        return loss.value, out

    def train(self, lat, dt, coarse_factor=8, npairs=1, lr=1e-3):
        losses = np.empty(npairs)
        self._check(lib().lat_corrector_train(self._h, lat._h, ctypes.c_double(dt), coarse_factor, npairs,
                                              ctypes.c_double(lr), _p(losses)), "train")
        return losses

    def step(self, lat, dt, coarse_factor=8, nsteps=1, threshold=1e-6):
        """Corrected coarse steps; returns dict(fallbacks, max_violation, force_evals)."""
        st = np.zeros(3)
        self._check(lib().lat_corrector_step(self._h, lat._h, ctypes.c_double(dt), coarse_factor, nsteps,
                                             ctypes.c_double(threshold), _p(st)), "step")
        return {"fallbacks": int(st[0]), "max_violation": float(st[1]), "force_evals": float(st[2])}

    def step_judged(self, lat, dt, coarse_factor=8, nsteps=1, threshold=1e-6, judge_every=4, err_tol=0.1):
        """Corrected coarse steps with the plain integrator as judge: every
        judge_every-th accepted step is also taken by the plain integrator,
        the errors of the corrected and of the leapfrog step against it are
        measured and the plain step kept; the corrector is switched off once
        a corrected step's error exceeds err_tol or the leapfrog step's."""
        st = np.zeros(7)
        self._check(lib().lat_corrector_step_judged(self._h, lat._h, ctypes.c_double(dt), coarse_factor, nsteps,
                                                    ctypes.c_double(threshold), judge_every,
                                                    ctypes.c_double(err_tol), _p(st)), "step_judged")
        return {"fallbacks": int(st[0]), "force_evals": float(st[1]), "judged": int(st[2]),
                "max_judged_error": float(st[3]), "switched_off_at": int(st[4]), "max_violation": float(st[5]),
                "max_judged_leapfrog_error": float(st[6])}
