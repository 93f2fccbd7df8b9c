"""GPU parity of the evolution path (splits, Yoshida compositions, calc_metric,
energies) against the CPU oracle on identical seeded initial conditions."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

PHI0_C, DPHI0_C = 1.0093430384226378929425913902459, -0.7137133070120812430962278466136


def two_field_ic(n, L, seed, mpl=2e5):
    """Seeded vacuum-like fluctuations (oracle sampler, Philox streams) on top of
    the homogeneous two-field background of hamiltonian_cosmic.f90:18-22."""
    m2eff = 1.0
    fld = np.empty((2, n, n, n))
    fldp = np.empty_like(fld)
    fld[0] = PHI0_C + O.sample(n, L, -0.25, m2eff, mpl, seed, 0)
    fldp[0] = DPHI0_C + O.sample(n, L, 0.25, m2eff, mpl, seed, 1)
    fld[1] = O.sample(n, L, -0.25, m2eff, mpl, seed, 2)
    fldp[1] = O.sample(n, L, 0.25, m2eff, mpl, seed, 3)
    return fld, fldp


def one_field_ic(n, L, seed, phi0=1.0, mpl=1e7 / 3.0):
    m2eff = 3.0 * phi0**2
    dphi0 = -phi0**3 / (3.0 * np.sqrt(phi0**4 / 12.0))
    fld = (phi0 + O.sample(n, L, -0.25, m2eff, mpl, seed, 0))[None]
    fldp = (dphi0 + O.sample(n, L, 0.25, m2eff, mpl, seed, 1))[None]
    return fld, fldp


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def run_pair(gpu, n, L, nfld, time_var, model, mp, lap, stencil, fld, fldp, plan):
    with gpu.Lattice(n, nfld=nfld, length=L, time_var=time_var, model=model, model_params=mp,
                     laplacian=lap, stencil=stencil) as lat:
        sim = O.Simulation(n, nfld, L, time_var, lap_kind=lap, stencil=stencil, model=model, mp=mp)
        lat.set_state(fld, fldp, 1.0, 0.0)
        sim.set_state(fld, fldp, 1.0, 0.0)
        lat.calc_metric()
        sim.calc_metric()
        y0, p0 = lat.scalars
        assert abs(p0 - sim.ysclp) <= 1e-13 * abs(sim.ysclp)
        for kind, args in plan:
            getattr(lat, kind)(*args)
            if kind == "symp":
                sim.symp(*args)
            elif kind == "Hamiltonian_Split":
                sim.split(*args)
            elif kind == "symp_o2step":
                sim.symp_o2step(*args)
        gf, gp, gy, gpy = lat.get_state()
        ge = lat.energies()
        oe = sim.energies()
        return (gf, gp, gy, gpy, ge), (sim.fld.copy(), sim.fldp.copy(), sim.yscl, sim.ysclp, oe)


CASES = {
    "conformal_phi4_spectral": dict(nfld=1, time_var=O.CONFORMAL, model=O.MODEL_PHI4, mp=(1.0,),
                                    lap=O.LAP_SPECTRAL, stencil=(-6, 1, 0, 0, 1)),
    "cosmic_twofield_spectral": dict(nfld=2, time_var=O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, 2.5e5),
                                     lap=O.LAP_SPECTRAL, stencil=(-6, 1, 0, 0, 1)),
    "cosmic_twofield_fd7": dict(nfld=2, time_var=O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, 2.5e5),
                                lap=O.LAP_STENCIL, stencil=(-6.0, 1.0, 0.0, 0.0, 1.0)),
    "conformal_phi4_fd27": dict(nfld=1, time_var=O.CONFORMAL, model=O.MODEL_PHI4, mp=(1.0,),
                                lap=O.LAP_STENCIL, stencil=(-128.0, 14.0, 3.0, 1.0, 30.0)),
    # the chi mass of the Discrete model (model.f90:35-55), m2psi = 3e3
    "cosmic_twofield_mchi_fd7": dict(nfld=2, time_var=O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, 2.5e5, 3.0e3),
                                     lap=O.LAP_STENCIL, stencil=(-6.0, 1.0, 0.0, 0.0, 1.0)),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_symp_matches_oracle(gpu, case, order):
    print("case", case, order, flush=True)
    c = CASES[case]
    n, L = 16, 10.0 if c["nfld"] == 2 else 20.0
    # two-field: m_chi,eff = g phi ~ 500 needs omega*dt well inside the symp stability region
    dt = 1e-3 if c["nfld"] == 2 else (L / n) / 40.0
    fld, fldp = (two_field_ic(n, L, 101) if c["nfld"] == 2 else one_field_ic(n, L, 37))
    g, o = run_pair(gpu, n, L, c["nfld"], c["time_var"], c["model"], c["mp"], c["lap"], c["stencil"],
                    fld, fldp, [("symp", (order, dt, 5)), ("symp", (order, dt, 3))])
    assert rel(g[0], o[0]) < 1e-12
    assert rel(g[1], o[1]) < 1e-12
    assert abs(g[2] - o[2]) < 1e-13 * abs(o[2])
    assert abs(g[3] - o[3]) < 1e-11 * abs(o[3])
    for k in ("rho", "ke", "pe", "ge", "resid", "a", "hubble"):
        assert abs(g[4][k] - o[4][k]) <= 1e-11 * abs(o[4][k]) + 1e-15, k


@pytest.mark.parametrize("case", list(CASES))
def test_single_splits_match_oracle(gpu, case):
    c = CASES[case]
    n, L = 16, 10.0
    fld, fldp = (two_field_ic(n, L, 5) if c["nfld"] == 2 else one_field_ic(n, L, 6))
    plan = [("Hamiltonian_Split", (0.01, 1)), ("Hamiltonian_Split", (0.02, 3)), ("Hamiltonian_Split", (0.01, 2)),
            ("symp_o2step", (0.01, 0.7, 0.3)), ("Hamiltonian_Split", (0.005, 3))]
    g, o = run_pair(gpu, n, L, c["nfld"], c["time_var"], c["model"], c["mp"], c["lap"], c["stencil"],
                    fld, fldp, plan)
    assert rel(g[0], o[0]) < 1e-13
    assert rel(g[1], o[1]) < 1e-13
    assert abs(g[2] - o[2]) < 1e-14 * abs(o[2])
    assert abs(g[3] - o[3]) < 1e-12 * abs(o[3])


def test_state_roundtrip_ghost_layout(gpu):
    """AoS ghosted fld(nfld,0:nx+1,...) in, periodic ghosts out (wrap_fields)."""
    n = 8
    rng = np.random.default_rng(0)
    fld = rng.standard_normal((2, n, n, n))
    fldp = rng.standard_normal((2, n, n, n))
    g = np.zeros((n + 2, n + 2, n + 2, 2))
    g[1:-1, 1:-1, 1:-1, :] = np.moveaxis(fld, 0, -1)
    O.wrap_fields(g)
    with gpu.Lattice(n, nfld=2, length=10.0) as lat:
        lat.set_state_aos(g, np.ascontiguousarray(np.moveaxis(fldp, 0, -1)), 1, 1.25, -0.5)
        aos, aosp, y, p = lat.get_state_aos(1)
    assert np.array_equal(aos, g)
    assert np.array_equal(aosp, np.moveaxis(fldp, 0, -1))
    assert (y, p) == (1.25, -0.5)
