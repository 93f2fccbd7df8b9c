"""FP64 parity of the evolution at the configurations' own sizes and run
lengths: the HIP path (through the C ABI) against the oracle from identical
seeded initial conditions, energies and output spectra after every output
interval (the reference's main loop, Pseudospec/openmp_dev/evolve_spectral.f90:93-97:
symp over the interval, then make_output).

Bars:
  * fields and momenta (full arrays), a, H, rho, KE, PE: <= 1e-11 relative;
  * quantities that FP64 itself resolves only coarsely -- the fluctuations
    (fields minus their means), the gradient energy, the Friedmann residual
    rho - 3H^2 (a difference of nearly equal numbers) and every spectrum bin --
    against the FP64 floor: within FLOOR_FAC times what a 1e-15 relative
    perturbation of the initial state moves them by in the oracle itself
    (tests/parity_util.py), and at most 1e-11 where that floor is smaller.
"""
import numpy as np
import pytest

import oracle as O
from parity_util import (E_KEYS, cfg0_ic, cfg1_ic, compare, fluct, perturb, rel, run_gpu, run_oracle)

pytestmark = pytest.mark.gpu

FLOOR_FAC = 100.0
STENCIL_7PT = (-6.0, 1.0, 0.0, 0.0, 1.0)
STENCIL_27PT = (-128.0, 14.0, 3.0, 1.0, 30.0)


def check_run(name, gpu_run, orc_run, floor_run):
    grec, gst = gpu_run
    orec, ost = orc_run
    prec_, pst = floor_run
    err = compare(grec, orec)
    flo = compare(prec_, orec)
    gf, gp, gy, gpy = gst
    of, op, oy, opy = ost
    pf, pp = pst[0], pst[1]
    line = {"fld": rel(gf, of), "fldp": rel(gp, op), "yscl": abs(gy - oy) / abs(oy),
            "fluct_fld": [rel(fluct(gf[l]), fluct(of[l])) for l in range(len(of))],
            "fluct_fld_floor": [rel(fluct(pf[l]), fluct(of[l])) for l in range(len(of))],
            "fluct_fldp": [rel(fluct(gp[l]), fluct(op[l])) for l in range(len(op))],
            "fluct_fldp_floor": [rel(fluct(pp[l]), fluct(op[l])) for l in range(len(op))],
            "energies_max": {k: max(e[k] for e in err["e"]) for k in E_KEYS if k != "resid"},
            "energies_floor": {k: max(e[k] for e in flo["e"]) for k in E_KEYS if k != "resid"},
            "resid_abs_max": max(err["resid_abs"]), "resid_abs_floor": max(flo["resid_abs"]),
            "resid_traj": [(g["e"]["resid"], o["e"]["resid"]) for g, o in zip(grec, orec)]}
    if err["s"]:
        line["spectra_max"] = max(max(s) for s in err["s"])
        line["spectra_floor"] = max(max(s) for s in flo["s"])
    print(name, line, flush=True)
    assert line["fld"] <= 1e-11 and line["fldp"] <= 1e-11 and line["yscl"] <= 1e-11, line
    for l in range(len(of)):
        assert line["fluct_fld"][l] <= max(FLOOR_FAC * line["fluct_fld_floor"][l], 1e-11), (l, line)
        assert line["fluct_fldp"][l] <= max(FLOOR_FAC * line["fluct_fldp_floor"][l], 1e-11), (l, line)
    for ec, ef in zip(err["e"], flo["e"]):
        for k in ("a", "hubble", "rho", "ke", "pe"):
            assert ec[k] <= 1e-11, (k, ec[k])
        assert ec["ge"] <= max(FLOOR_FAC * ef["ge"], 1e-11), (ec["ge"], ef["ge"])
    for ga, fa in zip(err["resid_abs"], flo["resid_abs"]):
        assert ga <= max(FLOOR_FAC * fa, 1e-15), (ga, fa)
    for sg, sf in zip(err["s"], flo["s"]):
        for a, b in zip(sg, sf):
            assert a <= max(FLOOR_FAC * b, 1e-11), (sg, sf)
    return line


def three_runs(gpu, n, L, nfld, time_var, model, mp, lap, stencil, fld, fldp, dt, nchunks, chunk, spectra=True):
    print("case", n, nfld, time_var, lap, stencil, dt, nchunks, chunk, flush=True)
    with gpu.Lattice(n, nfld=nfld, length=L, time_var=time_var, model=model, model_params=mp, laplacian=lap,
                     stencil=stencil) as lat:
        lat.set_state(fld, fldp, 1.0, 0.0)
        lat.calc_metric()
        g = run_gpu(lat, dt, nchunks, chunk, spectra)
    runs = [g]
    for eps in (0.0, 1e-15):
        sim = O.Simulation(n, nfld, L, time_var, lap_kind=lap, stencil=stencil, model=model, mp=mp)
        sim.set_state(perturb(fld, eps, 1), perturb(fldp, eps, 2), 1.0, 0.0)
        sim.calc_metric()
        runs.append(run_oracle(sim, dt, nchunks, chunk, spectra))
    return runs


def test_config0_64_conformal_phi4_1000_steps(gpu):
    """Config 0: 64^3, L = 20, conformal FRW, V = phi^4/4, spectral
    Laplacian, Yoshida-4, dt = dx/20, 1000 steps, energies and spectra every
    50 steps (20 outputs), residual trajectory overlaid."""
    n, L = 64, 20.0
    fld, fldp = cfg0_ic(n, L)
    runs = three_runs(gpu, n, L, 1, O.CONFORMAL, O.MODEL_PHI4, (1.0,), O.LAP_SPECTRAL, STENCIL_7PT, fld, fldp,
                      (L / n) / 20.0, 20, 50)
    check_run("config0", *runs)


def test_config1_256_spectral(gpu):
    """Config 1, spectral: 256^3 two-field preheating (g^2/m^2 = 2.5e5),
    cosmic FRW, L = 10, dt = dx/40, 60 Yoshida-4 steps, spectra of both
    fields every 20 steps."""
    n, L = 256, 10.0
    fld, fldp = cfg1_ic(n, L, 2024)
    runs = three_runs(gpu, n, L, 2, O.COSMIC, O.MODEL_TWOFIELD, (1.0, 2.5e5), O.LAP_SPECTRAL, STENCIL_7PT, fld,
                      fldp, (L / n) / 40.0, 3, 20)
    check_run("config1_spectral", *runs)


def test_config1_256_fd7(gpu):
    """Config 1, 7-point FD Laplacian: 256^3 (eight 32-plane chunks of the
    z-marching kick), 100 steps, spectra every 20 steps."""
    n, L = 256, 10.0
    fld, fldp = cfg1_ic(n, L, 2025)
    runs = three_runs(gpu, n, L, 2, O.COSMIC, O.MODEL_TWOFIELD, (1.0, 2.5e5), O.LAP_STENCIL, STENCIL_7PT, fld,
                      fldp, (L / n) / 40.0, 5, 20)
    check_run("config1_fd7", *runs)


@pytest.mark.parametrize("shape,stencil", [((48, 48, 48), STENCIL_7PT), ((96, 96, 96), STENCIL_7PT),
                                           ((64, 64, 64), STENCIL_7PT), ((32, 16, 48), STENCIL_7PT),
                                           ((48, 48, 48), STENCIL_27PT)],
                         ids=["48-7pt", "96-7pt", "64-7pt", "32x16x48-7pt", "48-27pt"])
def test_fd_chunked_z_march_and_grad2(gpu, shape, stencil):
    """The z-marching 7-point kick in 32-plane chunks: a partial last chunk
    (48), whole chunks (64, 96), a chunk start k0 > 0 reading its k0-1
    plane; and the general 27-point kernel.  dx as in config 1 (L = 10 n/256),
    two-field cosmic, 40 steps; then the GRAD2 gradient energy (grad_en,
    Discrete/Entropy/homlat.f90:88-115) against the oracle on the same state
    and against the Laplacian form (grad_en_wlap, :120-147)."""
    nx, ny, nz = shape
    L = 10.0 * nx / 256
    rng = np.random.default_rng(nz)
    fld = np.stack([PHI0 + 1e-6 * rng.standard_normal((nz, ny, nx)), 1e-6 * rng.standard_normal((nz, ny, nx))])
    fldp = np.stack([DPHI0 + 1e-5 * rng.standard_normal((nz, ny, nx)), 1e-5 * rng.standard_normal((nz, ny, nx))])
    print("case fd", shape, stencil, flush=True)
    dt = (L / nx) / 40.0
    with gpu.Lattice(shape, nfld=2, length=L, laplacian=gpu.LAP_STENCIL, stencil=stencil,
                     time_var=gpu.COSMIC, model=gpu.MODEL_TWOFIELD, model_params=(1.0, 2.5e5)) as lat:
        lat.set_state(fld, fldp, 1.0, 0.0)
        e0 = lat.energies()
        lat.calc_metric()
        g = run_gpu(lat, dt, 2, 20, spectra=False)
        eg = lat.energies()
    sim = O.Simulation(shape, 2, L, O.COSMIC, lap_kind=O.LAP_STENCIL, stencil=stencil, model=O.MODEL_TWOFIELD,
                       mp=(1.0, 2.5e5))
    sim.set_state(fld, fldp, 1.0, 0.0)
    o0 = sim.energies()
    sim.calc_metric()
    o = run_oracle(sim, dt, 2, 20, spectra=False)
    gf, gp = g[1][0], g[1][1]
    of, op = o[1][0], o[1][1]
    print("fd", shape, rel(gf, of), rel(gp, op), [rel(fluct(gf[l]), fluct(of[l])) for l in range(2)],
          e0["ge_grad"], o0["ge_grad"], e0["ge"], flush=True)
    assert rel(gf, of) <= 1e-12 and rel(gp, op) <= 1e-12
    for l in range(2):
        assert rel(fluct(gf[l]), fluct(of[l])) <= 1e-9
        assert rel(fluct(gp[l]), fluct(op[l])) <= 1e-9
    for k in ("a", "hubble", "rho", "ke", "pe"):
        assert abs(eg[k] - o[0][-1]["e"][k]) <= 1e-11 * abs(o[0][-1]["e"][k]), k
    # GRAD2 form on identical state: same differences, other summation order
    assert abs(e0["ge_grad"] - o0["ge_grad"]) <= 1e-12 * abs(o0["ge_grad"])
    # after the run: the evolved fluctuations agree to ~1e-10 (above), and so
    # do the squared differences
    assert abs(eg["ge_grad"] - o[0][-1]["e"]["ge_grad"]) <= 1e-8 * abs(o[0][-1]["e"]["ge_grad"])
    # summation by parts: GE_grad == GE_lap.  On phi = mean + 1e-6 noise the
    # Laplacian form loses ~1e-6 relative (the mean times the rounding of each
    # site's stencil sum), so the identity is checked on the fluctuations
    # alone, where both forms are well conditioned
    with gpu.Lattice(shape, nfld=2, length=L, laplacian=gpu.LAP_STENCIL, stencil=stencil,
                     time_var=gpu.COSMIC, model=gpu.MODEL_TWOFIELD, model_params=(1.0, 2.5e5)) as lat:
        lat.set_state(fld - fld.mean(axis=(1, 2, 3), keepdims=True), fldp, 1.0, 0.0)
        ez = lat.energies()
    assert abs(ez["ge_grad"] - ez["ge"]) <= 1e-12 * abs(ez["ge"]), (ez["ge_grad"], ez["ge"])


PHI0, DPHI0 = 1.0093430384226378929425913902459, -0.7137133070120812430962278466136


def test_grad2_energy_is_nan_for_spectral(gpu):
    with gpu.Lattice(16, nfld=2, length=10.0) as lat:
        rng = np.random.default_rng(1)
        lat.set_state(rng.standard_normal((2, 16, 16, 16)), rng.standard_normal((2, 16, 16, 16)), 1.0, 0.0)
        assert np.isnan(lat.energies()["ge_grad"])
