"""Mixed form (FP32 fluctuation fields and FFTs about FP64 homogeneous parts,
FP64 background and reductions) against the plain FP64 GPU form, from
identical initial conditions.  Tolerances are the stated errors of the mixed
form at these sizes/run lengths."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PHI0, DPHI0 = 1.0093430384226378929425913902459, -0.7137133070120812430962278466136


def run(gpu, prec, n, lap, steps, dt, chunks, seed=11):
    L, g2 = 10.0, 2.5e5
    lat = gpu.Lattice(n, nfld=2, length=L, precision=prec, laplacian=lap, stencil=gpu.STENCIL_7PT,
                      time_var=gpu.COSMIC, model=gpu.MODEL_TWOFIELD, model_params=(1.0, g2))
    lat.sample_field(0, 0, PHI0, -0.25, 1.0, 2e5, seed, 0)
    lat.sample_field(0, 1, DPHI0, 0.25, 1.0, 2e5, seed, 1)
    lat.sample_field(1, 0, 0.0, -0.25, g2 * PHI0**2, 2e5, seed, 2)
    lat.sample_field(1, 1, 0.0, 0.25, g2 * PHI0**2, 2e5, seed, 3)
    lat.set_scalars(1.0, 0.0)
    lat.calc_metric()
    e0 = lat.energies()
    spectra = []
    for _ in range(chunks):
        lat.symp4(dt, steps // chunks)
        spectra.append((lat.output_spectra(0), lat.output_spectra(1)))
    e1 = lat.energies()
    st = lat.get_state()
    lat.close()
    return e0, e1, spectra, st


def spec_err(a, b, floor=1e-6):
    """max relative error of S_phi and S_chi over k>0 bins that carry at least
    `floor` of the peak bin power (FP32 transforms have an absolute roundoff
    floor ~1e-7 of the rms mode amplitude; bins beyond the initial-condition
    cutoff carry ~1e-9 of the peak and are resolved only in FP64)."""
    err = 0.0
    for (a0, a1), (b0, b1) in zip(a, b):
        for x, y in ((a0[0], b0[0]), (a1[0], b1[0])):
            x, y = x[1:], y[1:]
            m = np.abs(y) >= floor * np.abs(y).max()
            err = max(err, float(np.max(np.abs(x[m] - y[m]) / np.abs(y[m]))))
    return err


@pytest.mark.parametrize("lap", ["spectral", "fd7"])
def test_mixed_tracks_fp64(gpu, lap):
    print("case mixed", lap, flush=True)
    n, steps, dt = 32, 200, 1e-3
    lp = gpu.LAP_SPECTRAL if lap == "spectral" else gpu.LAP_STENCIL
    e0d, e1d, sd, std_ = run(gpu, gpu.F64, n, lp, steps, dt, 4)
    e0m, e1m, sm, stm = run(gpu, gpu.MIXED, n, lp, steps, dt, 4)
    # initial state identical up to FP32 rounding of the fluctuations
    assert abs(e0m["rho"] - e0d["rho"]) <= 1e-9 * abs(e0d["rho"])
    assert abs(e1m["rho"] - e1d["rho"]) <= 1e-9 * abs(e1d["rho"])
    assert abs(e1m["a"] - e1d["a"]) <= 1e-11
    assert abs(e1m["resid"]) < 1e-7 and abs(e1d["resid"]) < 1e-7
    err3, err6 = spec_err(sm, sd, 1e-3), spec_err(sm, sd, 1e-6)
    print("mixed spectrum max rel err: floor 1e-3", err3, "floor 1e-6", err6, "all bins", spec_err(sm, sd, 0.0),
          flush=True)
    assert err3 < 1e-5
    assert err6 < 1e-4
    # bin 0 of the inflaton (N^3 times the FP64 mean plus the fluctuation sum,
    # the mixed form's FP32 k = 0 coefficient): S_phi, S_pi and Re S_pi_phi
    for (m0, _), (d0, _) in zip(sm, sd):
        for q in range(3):
            assert abs(m0[q][0] - d0[q][0]) <= 1e-8 * abs(d0[q][0]), (q, m0[q][0], d0[q][0])
    # fluctuation fields
    for l in range(2):
        df = stm[0][l] - std_[0][l]
        ref = std_[0][l] - std_[0][l].mean()
        assert np.linalg.norm(df) <= 1e-4 * np.linalg.norm(ref)


def test_mixed_bin0_of_zero_mean_field(gpu):
    """Bin 0 of chi, whose mean is ~0: S(0) = F(0)^2 with F(0) = N^3 <chi>
    taken from the FP32 transform's k = 0 coefficient (plus the FP64 mean).
    Its rounding is that of an FP32 sum of N^3 fluctuations, so
    |sqrt S_mixed(0) - sqrt S_f64(0)| <= C eps32 sqrt(N^3) rms(chi) with
    C = 4 sqrt(log2 N^3) (a random walk over the transform's passes), checked
    on the identical initial state of both forms for S_chi and S_pi_chi."""
    print("case mixed bin0 chi", flush=True)
    n = 32
    eps32 = 2.0 ** -24
    out = {}
    for prec in (gpu.F64, gpu.MIXED):
        lat = gpu.Lattice(n, nfld=2, length=10.0, precision=prec, time_var=gpu.COSMIC,
                          model=gpu.MODEL_TWOFIELD, model_params=(1.0, 2.5e5))
        lat.sample_field(0, 0, PHI0, -0.25, 1.0, 2e5, 11, 0)
        lat.sample_field(0, 1, DPHI0, 0.25, 1.0, 2e5, 11, 1)
        lat.sample_field(1, 0, 0.0, -0.25, 2.5e5 * PHI0 ** 2, 2e5, 11, 2)
        lat.sample_field(1, 1, 0.0, 0.25, 2.5e5 * PHI0 ** 2, 2e5, 11, 3)
        out[prec] = (lat.output_spectra(1), lat.get_state())
        lat.close()
    (sd, st), (sm, _) = out[gpu.F64], out[gpu.MIXED]
    c = 4.0 * np.sqrt(np.log2(n ** 3))
    for q, arr in ((0, st[0][1]), (1, st[1][1])):
        bound = c * eps32 * np.sqrt(n ** 3) * np.sqrt(np.mean(arr ** 2))
        d = abs(np.sqrt(sm[q][0]) - np.sqrt(sd[q][0]))
        print("chi bin0", q, np.sqrt(sd[q][0]), d, bound, flush=True)
        assert d <= bound, (q, d, bound)


def test_mixed_tracks_fp64_256_1000_steps(gpu):
    """Config-1 physics at 256^3 over 1000 Yoshida-4 steps (dt = dx/40), both
    forms from identical device-sampled initial conditions, spectra every
    100 steps.  Stated: fluctuation fields, spectra (bins >= 1e-3 and >= 1e-6
    of the peak, and all bins), energy and Friedmann residual of the mixed
    form against the FP64 form."""
    print("case mixed 256 x 1000", flush=True)
    n, steps, chunks = 256, 1000, 10
    dt = (10.0 / n) / 40.0
    e0d, e1d, sd, std_ = run(gpu, gpu.F64, n, gpu.LAP_SPECTRAL, steps, dt, chunks, seed=21)
    e0m, e1m, sm, stm = run(gpu, gpu.MIXED, n, gpu.LAP_SPECTRAL, steps, dt, chunks, seed=21)
    fl = [float(np.linalg.norm(stm[0][l] - std_[0][l]) / np.linalg.norm(std_[0][l] - std_[0][l].mean()))
          for l in range(2)]
    flp = [float(np.linalg.norm(stm[1][l] - std_[1][l]) / np.linalg.norm(std_[1][l] - std_[1][l].mean()))
           for l in range(2)]
    err3, err6, errall = spec_err(sm, sd, 1e-3), spec_err(sm, sd, 1e-6), spec_err(sm, sd, 0.0)
    erho = abs(e1m["rho"] - e1d["rho"]) / abs(e1d["rho"])
    dres = abs(e1m["resid"] - e1d["resid"])
    print("mixed 256^3 x 1000:", {"fluct_fld": fl, "fluct_fldp": flp, "spec_ge_1e-3": err3, "spec_ge_1e-6": err6,
                                  "spec_all": errall, "rho": erho, "resid_mixed": e1m["resid"],
                                  "resid_f64": e1d["resid"]}, flush=True)
    # measured on MI355X: fluctuations 9.0e-6 / 2.7e-6 (fields), 8.6e-6 /
    # 2.5e-6 (momenta); spectra 1.2e-6 (>= 1e-3 of peak), 4.9e-6 (>= 1e-6),
    # 2.0e-2 (all bins); rho 6.0e-11; residual 9e-11 apart
    assert max(fl) < 5e-5 and max(flp) < 5e-5
    assert err3 < 1e-5
    assert err6 < 3e-5
    assert erho < 1e-9
    assert dres < 1e-9


def test_mixed_with_chi_mass_tracks_oracle(gpu):
    """The chi mass term of the Discrete model (V += m2chi chi^2 / 2) in the
    mixed form's difference forms of V and V' (fluctuations about the FP64
    homogeneous parts) and in the FP64 form, against the oracle."""
    import oracle as O

    print("case mixed chi mass", flush=True)
    n, L, dt, nsteps = 32, 10.0, 1e-3, 20
    mp = (1.0, 2.5e5, 4.0e3)
    rng = np.random.default_rng(9)
    fld = np.stack([PHI0 + 1e-5 * rng.standard_normal((n, n, n)), 2e-3 + 1e-5 * rng.standard_normal((n, n, n))])
    fldp = np.stack([DPHI0 + 1e-4 * rng.standard_normal((n, n, n)), 1e-4 * rng.standard_normal((n, n, n))])
    sim = O.Simulation(n, 2, L, O.COSMIC, model=O.MODEL_TWOFIELD, mp=mp)
    sim.set_state(fld, fldp, 1.0, 0.0)
    sim.calc_metric()
    sim.symp(4, dt, nsteps)
    # errors relative to the fluctuations (1e-5 .. 1e-4 on means of order 1):
    # FP64 resolves them to ~1e-16 / 1e-5, the mixed form to FP32 rounding
    for prec, tol in ((gpu.F64, 1e-9), (gpu.MIXED, 1e-5)):
        with gpu.Lattice(n, nfld=2, length=L, precision=prec, model_params=mp) as lat:
            lat.set_state(fld, fldp, 1.0, 0.0)
            lat.calc_metric()
            lat.symp4(dt, nsteps)
            f, p, y, py = lat.get_state()
        if prec == gpu.F64:
            assert np.linalg.norm(f - sim.fld) <= 1e-12 * np.linalg.norm(sim.fld)
            assert np.linalg.norm(p - sim.fldp) <= 1e-12 * np.linalg.norm(sim.fldp)
        for l in range(2):
            for a, b in ((f[l], sim.fld[l]), (p[l], sim.fldp[l])):
                e = np.linalg.norm(a - b) / np.linalg.norm(b - b.mean())
                assert e < tol, (prec, l, e)
        assert abs(y - sim.yscl) < 1e-13


def test_mixed_state_roundtrip(gpu):
    n = 16
    rng = np.random.default_rng(3)
    fld = np.empty((2, n, n, n))
    fld[0] = PHI0 + 1e-6 * rng.standard_normal((n, n, n))
    fld[1] = 1e-6 * rng.standard_normal((n, n, n))
    fldp = DPHI0 + 1e-6 * rng.standard_normal((2, n, n, n))
    with gpu.Lattice(n, nfld=2, length=10.0, precision=gpu.MIXED) as lat:
        lat.set_state(fld, fldp, 1.0, -0.5)
        f, p, y, py = lat.get_state()
    for l in range(2):  # fluctuations ~1e-6 stored in FP32: ~1e-13 absolute
        assert np.abs(f[l] - fld[l]).max() <= 1e-12
        assert np.abs(p[l] - fldp[l]).max() <= 1e-12


def test_mixed_tracks_fp64_at_1024(gpu):
    """The 1024^3 launch shapes (fused z pass, LDS-streamed FP32 y passes,
    one-block-per-tile x sweeps) against the FP64 form: energies and spectra
    only (no host copy of the 8 GiB fields)."""
    print("case mixed 1024", flush=True)
    n, steps, dt = 1024, 4, (10.0 / 1024) / 40
    out = {}
    for prec in (gpu.F64, gpu.MIXED):
        L, g2 = 10.0, 2.5e5
        with gpu.Lattice(n, nfld=2, length=L, precision=prec, time_var=gpu.COSMIC, model=gpu.MODEL_TWOFIELD,
                         model_params=(1.0, g2)) as lat:
            lat.sample_field(0, 0, PHI0, -0.25, 1.0, 2e5, 5, 0)
            lat.sample_field(0, 1, DPHI0, 0.25, 1.0, 2e5, 5, 1)
            lat.sample_field(1, 0, 0.0, -0.25, g2 * PHI0**2, 2e5, 5, 2)
            lat.sample_field(1, 1, 0.0, 0.25, g2 * PHI0**2, 2e5, 5, 3)
            lat.set_scalars(1.0, 0.0)
            lat.calc_metric()
            lat.symp4(dt, steps)
            out[prec] = (lat.energies(), [(lat.output_spectra(0), lat.output_spectra(1))])
    (ed, sd), (em, sm) = out[gpu.F64], out[gpu.MIXED]
    assert abs(em["rho"] - ed["rho"]) <= 1e-8 * abs(ed["rho"])
    assert abs(em["a"] - ed["a"]) <= 1e-12
    err3 = spec_err(sm, sd, 1e-3)
    print("1024 mixed spectrum max rel err (floor 1e-3)", err3, flush=True)
    assert err3 < 1e-5
