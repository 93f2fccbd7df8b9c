"""GPU parity of the metric-perturbation (gravitational wave) extension
against the CPU oracle: central-difference source S_ij, the u_ij drift and
kick inside the Yoshida compositions (spectral and FD Laplacians, both time
variables), state transfer in the hij(2,3,n,n,n) layout, and the TT-projected
GW spectra.  The reference holds only stubs for this part (grav-waves/), so
these tests pin the GPU to the oracle, and the oracle to closed forms
(tests/test_oracle_pins.py)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def fields_ic(n, L, nfld, seed, phi0=(1.0, 0.0), mpl=20.0):
    """Seeded fluctuations of a few percent on homogeneous values."""
    fld = np.empty((nfld, n, n, n))
    fldp = np.empty_like(fld)
    for l in range(nfld):
        fld[l] = phi0[l] + O.sample(n, L, -0.25, 1.0, mpl, seed, 2 * l)
        fldp[l] = O.sample(n, L, 0.25, 1.0, mpl, seed, 2 * l + 1)
    return fld, fldp


def gw_ic(n, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    u = amp * rng.standard_normal((6, n, n, n))
    pu = amp * rng.standard_normal((6, n, n, n))
    return u, pu


CASES = {
    "conformal_phi4_spectral": dict(nfld=1, time_var=O.CONFORMAL, model=O.MODEL_PHI4, mp=(1.0,),
                                    lap=O.LAP_SPECTRAL, stencil=(-6.0, 1.0, 0.0, 0.0, 1.0)),
    "cosmic_twofield_spectral": dict(nfld=2, time_var=O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, 100.0),
                                     lap=O.LAP_SPECTRAL, stencil=(-6.0, 1.0, 0.0, 0.0, 1.0)),
    "cosmic_twofield_fd7": dict(nfld=2, time_var=O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, 100.0),
                                lap=O.LAP_STENCIL, stencil=(-6.0, 1.0, 0.0, 0.0, 1.0)),
    "conformal_phi4_fd27": dict(nfld=1, time_var=O.CONFORMAL, model=O.MODEL_PHI4, mp=(1.0,),
                                lap=O.LAP_STENCIL, stencil=(-128.0, 14.0, 3.0, 1.0, 30.0)),
}


def make_pair(gpu, c, n, L, gw_mpl, precision=None):
    kw = dict(nfld=c["nfld"], length=L, time_var=c["time_var"], model=c["model"], model_params=c["mp"],
              laplacian=c["lap"], stencil=c["stencil"], gw=True, gw_mpl=gw_mpl)
    if precision is not None:
        kw["precision"] = precision
    lat = gpu.Lattice(n, **kw)
    sim = O.Simulation(n, c["nfld"], L, c["time_var"], lap_kind=c["lap"], stencil=c["stencil"], model=c["model"],
                       mp=c["mp"], gw=True, gw_mpl=gw_mpl)
    return lat, sim


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_gw_source_matches_oracle(gpu, precision):
    print("case gw source", precision, flush=True)
    n, L = 16, 10.0
    fld, fldp = fields_ic(n, L, 2, 5)
    prec = gpu.F64 if precision == "f64" else gpu.MIXED
    with gpu.Lattice(n, nfld=2, length=L, precision=prec, gw=True) as lat:
        lat.set_state(fld, fldp, 1.0, 0.0)
        S = lat.gw_source()
    ref = O.gw_source_fd(fld, L / n)
    # mixed: gradients of the FP32 fluctuations (about 1e-7 relative each)
    tol = 1e-13 if precision == "f64" else 2e-6
    for c in range(6):
        assert rel(S[c], ref[c]) < tol, c


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_gw_state_roundtrip(gpu, precision):
    print("case gw roundtrip", precision, flush=True)
    n = 8
    u, pu = gw_ic(n, 3)
    prec = gpu.F64 if precision == "f64" else gpu.MIXED
    with gpu.Lattice(n, nfld=2, length=5.0, precision=prec, gw=True) as lat:
        lat.set_gw(u, pu)
        u2, p2 = lat.get_gw()
        # the reference layout hij(2,3,...): (1,b) diagonal, (2,b) = 12, 13, 23
        aos = np.empty((n, n, n, 3, 2))
        aos[..., 0, 0], aos[..., 1, 0], aos[..., 2, 0] = u[0], u[3], u[5]
        aos[..., 0, 1], aos[..., 1, 1], aos[..., 2, 1] = u[1], u[2], u[4]
        lat.set_gw_aos(aos, aos)
        u3, p3 = lat.get_gw()
    tol = 0.0 if precision == "f64" else 1e-7
    assert np.abs(u2 - u).max() <= tol * np.abs(u).max()
    assert np.abs(p2 - pu).max() <= tol * np.abs(pu).max()
    assert np.abs(u3 - u).max() <= tol * np.abs(u).max()
    assert np.abs(p3 - u).max() <= tol * np.abs(u).max()


def test_gw_needs_gw_context(gpu):
    with gpu.Lattice(8, nfld=2, length=5.0) as lat:
        with pytest.raises(gpu.LatticeError):
            lat.gw_spectrum()


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("order", [2, 4])
def test_gw_evolution_matches_oracle(gpu, case, order):
    print("case gw evolution", case, order, flush=True)
    c = CASES[case]
    n, L, dt, nsteps = 16, 10.0, 0.01, 6
    fld, fldp = fields_ic(n, L, c["nfld"], 11)
    u, pu = gw_ic(n, 12)
    lat, sim = make_pair(gpu, c, n, L, gw_mpl=0.5)
    with lat:
        lat.set_state(fld, fldp, 1.0, 0.0)
        sim.set_state(fld, fldp, 1.0, 0.0)
        lat.set_gw(u, pu)
        sim.gwu[...] = u
        sim.gwp[...] = pu
        lat.calc_metric()
        sim.calc_metric()
        lat.symp(order, dt, nsteps)
        sim.symp(order, dt, nsteps)
        gf, gp, gy, gpy = lat.get_state()
        gu, gpu_ = lat.get_gw()
    assert abs(gy - sim.yscl) <= 1e-13 * abs(sim.yscl)
    assert rel(gf, sim.fld) < 1e-12
    assert rel(gp, sim.fldp) < 1e-11
    # the sourced part must be visible: compare with the unsourced evolution size
    assert rel(gpu_, pu) > 1e-3
    assert rel(gu, sim.gwu) < 1e-11
    assert rel(gpu_, sim.gwp) < 1e-11


@pytest.mark.parametrize("lap", ["spectral", "fd7"])
def test_gw_split_pieces_match_oracle(gpu, lap):
    """Single split calls (drift, grav, kick) and symp_o2step with u_ij."""
    print("case gw splits", lap, flush=True)
    c = CASES["cosmic_twofield_" + lap]
    n, L = 16, 10.0
    fld, fldp = fields_ic(n, L, 2, 21)
    u, pu = gw_ic(n, 22)
    lat, sim = make_pair(gpu, c, n, L, gw_mpl=1.0)
    with lat:
        lat.set_state(fld, fldp, 1.0, 0.0)
        sim.set_state(fld, fldp, 1.0, 0.0)
        lat.set_gw(u, pu)
        sim.gwu[...] = u
        sim.gwp[...] = pu
        lat.calc_metric()
        sim.calc_metric()
        for h, ind in [(0.02, 2), (0.01, 3), (0.015, 1), (0.02, 2), (0.02, 2), (0.01, 3)]:
            lat.Hamiltonian_Split(h, ind)
            sim.split(h, ind)
        lat.symp_o2step(0.01, 1.0, 0.5)
        sim.symp_o2step(0.01, 1.0, 0.5)
        gu, gp = lat.get_gw()
    assert rel(gu, sim.gwu) < 1e-12
    assert rel(gp, sim.gwp) < 1e-12


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 16, 8), (48, 32, 16)], ids=lambda s: "x".join(map(str, s)))
def test_gw_spectrum_matches_oracle(gpu, shape):
    print("case gw spectrum", shape, flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(nx + 3 * ny)
    u = rng.standard_normal((6, nz, ny, nx))
    pu = rng.standard_normal((6, nz, ny, nx))
    with gpu.Lattice(shape, nfld=2, length=7.0, gw=True) as lat:
        lat.set_gw(u, pu)
        S = lat.gw_spectrum()
        S2 = lat.gw_spectrum()
    for ch, arr in enumerate((pu, u)):
        Uk = np.stack([O.r2c_3d(arr[q]) for q in range(6)])
        ref = O.gw_spectrum(O.tt_project(Uk, nx), nx)
        assert np.allclose(S[ch], ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max()), ch
    assert np.array_equal(S, S2)


def test_gw_mixed_form_tracks_f64(gpu):
    """FP32 u_ij / fields with FP64 background against the FP64 form."""
    print("case gw mixed", flush=True)
    c = CASES["cosmic_twofield_spectral"]
    n, L, dt, nsteps = 32, 10.0, 0.01, 10
    fld, fldp = fields_ic(n, L, 2, 31)
    u, pu = gw_ic(n, 32, amp=1e-4)
    out = {}
    for name, prec in (("f64", gpu.F64), ("mixed", gpu.MIXED)):
        kw = dict(nfld=2, length=L, time_var=c["time_var"], model=c["model"], model_params=c["mp"],
                  laplacian=c["lap"], gw=True, gw_mpl=0.5, precision=prec)
        with gpu.Lattice(n, **kw) as lat:
            lat.set_state(fld, fldp, 1.0, 0.0)
            lat.set_gw(u, pu)
            lat.calc_metric()
            lat.symp(4, dt, nsteps)
            out[name] = lat.get_gw() + (lat.gw_spectrum(),)
    assert rel(out["mixed"][0], out["f64"][0]) < 1e-5
    assert rel(out["mixed"][1], out["f64"][1]) < 1e-5
    s_m, s_d = out["mixed"][2], out["f64"][2]
    m = s_d > 1e-6 * s_d.max()
    assert np.max(np.abs(s_m[m] - s_d[m]) / s_d[m]) < 1e-4


def test_gw_mixed_form_tracks_f64_at_128(gpu):
    """The headline physics with the six u_ij at 128^3 (u = pi_u = 0 at the
    start, sourced by the vacuum fluctuations of both fields), 100 Yoshida-4
    steps of dt = dx/40 in both forms from identical device-sampled initial
    conditions: GW spectra of pi_u and u (bins >= 1e-6 of the peak) and the
    u_ij fields of the mixed form against FP64."""
    print("case gw mixed 128", flush=True)
    n, L, g2 = 128, 10.0, 2.5e5
    phi0, dphi0 = 1.0093430384226378929425913902459, -0.7137133070120812430962278466136
    dt = (L / n) / 40.0
    out = {}
    for name, prec in (("f64", gpu.F64), ("mixed", gpu.MIXED)):
        with gpu.Lattice(n, nfld=2, length=L, precision=prec, time_var=gpu.COSMIC, model=gpu.MODEL_TWOFIELD,
                         model_params=(1.0, g2), gw=True, gw_mpl=1.0) as lat:
            lat.sample_field(0, 0, phi0, -0.25, 1.0, 2e5, 8, 0)
            lat.sample_field(0, 1, dphi0, 0.25, 1.0, 2e5, 8, 1)
            lat.sample_field(1, 0, 0.0, -0.25, g2 * phi0 ** 2, 2e5, 8, 2)
            lat.sample_field(1, 1, 0.0, 0.25, g2 * phi0 ** 2, 2e5, 8, 3)
            lat.set_scalars(1.0, 0.0)
            lat.calc_metric()
            lat.symp4(dt, 100)
            out[name] = lat.get_gw() + (lat.gw_spectrum(),)
    errs = []
    for ch in (0, 1):
        s_m, s_d = out["mixed"][2][ch][1:], out["f64"][2][ch][1:]
        m = s_d > 1e-6 * s_d.max()
        errs.append(float(np.max(np.abs(s_m[m] - s_d[m]) / s_d[m])))
    fu = [rel(out["mixed"][k], out["f64"][k]) for k in (0, 1)]
    print("gw mixed 128: spectra (pi_u, u)", errs, "u, pi_u fields", fu, flush=True)
    assert max(errs) < 1e-4
    assert max(fu) < 1e-4
