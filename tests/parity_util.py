"""Helpers of the long-run parity tests: identical seeded initial conditions
for the HIP path and the oracle, chunked Yoshida runs with energies and
output spectra after every chunk (the reference's main loop: symp4 over an
output interval, then make_output, Pseudospec/openmp_dev/evolve_spectral.f90:93-97),
and error summaries.

The FP64 floor: fields are a homogeneous part of order 1 plus fluctuations
of order 1e-7 .. 1e-5, so FP64 resolves the fluctuations (and every k > 0
spectrum bin) only to ~1e-16 |mean| / |fluctuation|.  The oracle run from
the same initial state perturbed by 1e-15 relative noise shows how far the
reference's own arithmetic can move each quantity; parity of quantities
that roundoff alone moves by more than 1e-11 is stated against that floor.
"""
import numpy as np

import oracle as O

PHI0_C, DPHI0_C = 1.0093430384226378929425913902459, -0.7137133070120812430962278466136
E_KEYS = ("a", "hubble", "rho", "ke", "pe", "ge", "resid")


def cfg0_ic(n, L, seed=37):
    """Config 0: lambda phi^4 in program units (V = phi^4/4), phi0 = 1, the
    slow-roll dphi0 and vacuum fluctuations at m2eff = 3 phi0^2 with the
    openmp_dev mpl = 1e7/3 (hamiltonian_conformal.f90:9, evolve_spectral.f90:289-300)."""
    phi0, mpl = 1.0, 1e7 / 3.0
    m2eff = 3.0 * phi0 ** 2
    dphi0 = -phi0 ** 3 / (3.0 * np.sqrt(phi0 ** 4 / 12.0))
    fld = (phi0 + O.sample(n, L, -0.25, m2eff, mpl, seed, 0))[None]
    fldp = (dphi0 + O.sample(n, L, 0.25, m2eff, mpl, seed, 1))[None]
    return fld, fldp


def cfg1_ic(n, L, seed, g2=2.5e5, mpl=2e5):
    """Two-field preheating background (hamiltonian_cosmic.f90:18-22) plus
    vacuum fluctuations of both fields at their effective masses."""
    m2 = [1.0, g2 * PHI0_C ** 2]
    fld = np.stack([PHI0_C + O.sample(n, L, -0.25, m2[0], mpl, seed, 0),
                    O.sample(n, L, -0.25, m2[1], mpl, seed, 2)])
    fldp = np.stack([DPHI0_C + O.sample(n, L, 0.25, m2[0], mpl, seed, 1),
                     O.sample(n, L, 0.25, m2[1], mpl, seed, 3)])
    return fld, fldp


def perturb(a, eps, seed):
    rng = np.random.default_rng(seed)
    return a * (1.0 + eps * rng.standard_normal(a.shape))


def oracle_spectra(sim):
    """make_output's spectra of every field (evolve_spectral.f90:531-538):
    S_phi, S_pi and crossspec_3d(pi^, phi^), as lat_output_spectra returns them."""
    out = []
    for l in range(sim.nfld):
        Fphi, Fpi = O.r2c_3d(sim.fld[l]), O.r2c_3d(sim.fldp[l])
        cre, cim = O.crossspec_3d(Fpi, Fphi, sim.nx)
        out.append(np.stack([O.spectrum_3d(Fphi, sim.nx), O.spectrum_3d(Fpi, sim.nx), cre, cim]))
    return out


def run_oracle(sim, dt, nchunks, chunk, spectra=True):
    rec = []
    for _ in range(nchunks):
        sim.symp(4, dt, chunk)
        rec.append({"e": sim.energies(), "s": oracle_spectra(sim) if spectra else None})
    return rec, (sim.fld.copy(), sim.fldp.copy(), sim.yscl, sim.ysclp)


def run_gpu(lat, dt, nchunks, chunk, spectra=True):
    rec = []
    for _ in range(nchunks):
        lat.symp4(dt, chunk)
        rec.append({"e": lat.energies(),
                    "s": [lat.output_spectra(l) for l in range(lat.nfld)] if spectra else None})
    f, p, y, py = lat.get_state()
    return rec, (f, p, y, py)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def fluct(a):
    return a - a.mean(axis=tuple(range(1, a.ndim)), keepdims=True)


def spec_err(a, b):
    """Per output and field: max over bins of |a-b|/|b| for the two power
    spectra, L2-relative for the cross spectrum (Re, Im together)."""
    errs = []
    for fa, fb in zip(a, b):
        for ch in (0, 1):
            m = np.abs(fb[ch]) > 0
            errs.append(float(np.max(np.abs(fa[ch][m] - fb[ch][m]) / np.abs(fb[ch][m]))))
        errs.append(rel(fa[2:], fb[2:]))
    return errs


def compare(rec_a, rec_b):
    """Errors of run a against run b after every chunk."""
    out = {"e": [], "resid_abs": [], "s": []}
    for ra, rb in zip(rec_a, rec_b):
        out["e"].append({k: abs(ra["e"][k] - rb["e"][k]) / abs(rb["e"][k]) for k in E_KEYS if k != "resid"})
        out["resid_abs"].append(abs(ra["e"]["resid"] - rb["e"]["resid"]))
        if ra["s"] is not None:
            out["s"].append(spec_err(ra["s"], rb["s"]))
    return out
