"""CPU pins of the oracle (oracle/liblatoracle.so) against what the
mathematics and independent libraries fix: numpy.fft goldens, closed forms,
brute-force sums, the Hamiltonian ODEs solved by scipy, convergence orders,
reversibility and invariants.  None of these retypes the oracle's formula."""
import os

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz")


# ---------------------------------------------------------------- FFT

def test_fft_matches_numpy_goldens():
    g = np.load(GOLDEN)
    keys = sorted(k[3:] for k in g.files if k.startswith("in_"))
    assert keys
    for key in keys:
        nx = int(key.split("x")[0])
        f = g[f"in_{key}"]
        F = O.r2c_3d(f)
        ref = g[f"r2c_{key}"]
        assert np.abs(F - ref).max() <= 1e-13 * np.abs(ref).max(), key
        G = g[f"cin_{key}"]
        out = O.c2r_3d(G, nx)
        ref2 = g[f"c2r_{key}"]
        assert np.abs(out - ref2).max() <= 1e-13 * np.abs(ref2).max(), key


@pytest.mark.parametrize("n", [2, 3, 5, 7, 12, 16, 30, 64, 96])
def test_fft_1d_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(j, j) / n) @ x
    assert np.abs(O.fft_c2c_1d(x, -1) - dft).max() < 1e-12 * np.abs(dft).max()
    idft = np.exp(2j * np.pi * np.outer(j, j) / n) @ x
    assert np.abs(O.fft_c2c_1d(x, +1) - idft).max() < 1e-12 * np.abs(idft).max()


def test_roundtrip_and_parseval():
    rng = np.random.default_rng(1)
    f = rng.standard_normal((12, 10, 16))
    F = O.r2c_3d(f)
    assert np.abs(O.c2r_3d(F, 16) / f.size - f).max() < 1e-14
    full = np.fft.fftn(f)
    assert abs(np.sum(np.abs(full) ** 2) / f.size - np.sum(f ** 2)) < 1e-10 * np.sum(f ** 2)


# ---------------------------------------------------------------- Laplacians

def plane(n, L, kv, phase=0.3):
    x = np.arange(n) * L / n
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    dk = 2 * np.pi / L
    return np.cos(dk * (kv[0] * x3 + kv[1] * y3 + kv[2] * z3) + phase)


@pytest.mark.parametrize("kv", [(1, 0, 0), (3, -2, 5), (0, 8, 0), (8, 8, 8), (-7, 1, 2)])
def test_spectral_laplacian_plane_waves(kv):
    n, L = 16, 7.0
    dk = 2 * np.pi / L
    f = plane(n, L, kv, phase=0.0 if 8 in kv else 0.3)  # Nyquist modes: cos only survives sampling
    lap = O.laplacian_3d(f, dk)
    expect = -dk**2 * (kv[0] ** 2 + kv[1] ** 2 + kv[2] ** 2) * f
    assert np.abs(lap - expect).max() < 1e-11 * max(1.0, np.abs(expect).max())


def test_spectral_laplacian_brute_force_dft():
    n, dk = 4, 0.9
    rng = np.random.default_rng(2)
    f = rng.standard_normal((n, n, n))
    idx = np.arange(n)
    kk = np.where(idx <= n // 2, idx, n - idx)  # reference magnitude convention, Nyquist kept
    E = np.exp(-2j * np.pi * np.outer(idx, idx) / n)
    F = np.einsum("ai,bj,ck,ijk->abc", E, E, E, f.astype(complex))
    k2 = kk[:, None, None] ** 2 + kk[None, :, None] ** 2 + kk[None, None, :] ** 2
    G = -dk**2 * k2 * F
    back = np.einsum("ai,bj,ck,ijk->abc", E.conj(), E.conj(), E.conj(), G).real / n**3
    assert np.abs(O.laplacian_3d(f, dk) - back).max() < 1e-12 * np.abs(back).max()


def test_gradient_squared_plane_wave():
    n, L = 16, 9.0
    dk = 2 * np.pi / L
    kv = (2, -3, 1)
    f = plane(n, L, kv)
    g2 = O.gradient_squared_3d(f, dk)
    x = np.arange(n) * L / n
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    expect = dk**2 * (4 + 9 + 1) * np.sin(dk * (2 * x3 - 3 * y3 + z3) + 0.3) ** 2
    assert np.abs(g2 - expect).max() < 1e-11


@pytest.mark.parametrize("coef", [(-6.0, 1.0, 0.0, 0.0, 1.0), (-128.0, 14.0, 3.0, 1.0, 30.0),
                                  (-24.0, 2.0, 1.0, 0.0, 6.0)])
def test_stencil_eigenvalue(coef):
    n, L = 12, 5.0
    dx = L / n
    kv = (1, 4, -2)
    f = plane(n, L, kv)
    th = [2 * np.pi * k / n for k in kv]
    c0, c1, c2, c3, cc = coef
    cs = [np.cos(t) for t in th]
    lam = (c0 + 2 * c1 * sum(cs) + 4 * c2 * (cs[0] * cs[1] + cs[0] * cs[2] + cs[1] * cs[2])
           + 8 * c3 * cs[0] * cs[1] * cs[2]) / (cc * dx**2)
    assert c0 + 6 * c1 + 12 * c2 + 8 * c3 == 0.0  # consistency: constants are annihilated
    out = O.stencil_laplacian(f, coef, dx)
    assert np.abs(out - lam * f).max() < 1e-11 * abs(lam)


def test_grad2_equals_lap_form_on_periodic_lattice():
    """homlat.f90:90-147: GE from squared differences == -phi*lap (summation by parts)."""
    rng = np.random.default_rng(4)
    f = rng.standard_normal((8, 8, 8))
    for coef in [(-6.0, 1.0, 0.0, 0.0, 1.0), (-128.0, 14.0, 3.0, 1.0, 30.0)]:
        g = O.stencil_grad2_sum(f, coef) / (4 * coef[4])
        lap = O.stencil_laplacian(f, coef, 1.0)
        assert abs(g - (-0.5 * np.sum(f * lap))) < 1e-10 * abs(g)


def test_wrap_fields_periodic():
    rng = np.random.default_rng(5)
    n, nf = 6, 2
    inner = rng.standard_normal((n, n, n, nf))
    g = np.zeros((n + 2, n + 2, n + 2, nf))
    g[1:-1, 1:-1, 1:-1] = inner
    O.wrap_fields(g)
    idx = (np.arange(n + 2) - 1) % n
    assert np.array_equal(g, inner[idx][:, idx][:, :, idx])


# ---------------------------------------------------------------- spectra

def brute_spectrum(f1, f2=None, n6=False, counts=False):
    """Binning over the full complex spectrum (numpy fftn, every k, no mirror)."""
    nz, ny, nx = f1.shape
    F1 = np.fft.fftn(f1)
    F2 = F1 if f2 is None else np.fft.fftn(f2)
    a = F1 * np.conj(F2)
    ii = np.minimum(np.arange(nx), nx - np.arange(nx))
    jj = np.minimum(np.arange(ny), ny - np.arange(ny))
    kk = np.minimum(np.arange(nz), nz - np.arange(nz))
    p = np.sqrt(kk[:, None, None] ** 2 + jj[None, :, None] ** 2 + ii[None, None, :] ** 2).ravel()
    ns = int(np.floor(np.sqrt((nx // 2 + 1) ** 2 + (ny // 2 + 1) ** 2 + (nz // 2 + 1) ** 2 + 1)))
    if counts:
        ns = int(1.7320508075688772935274463415059 * (nx // 2) + 2.0)
    l = np.floor(p).astype(int)
    S = np.zeros(ns, complex)
    W = np.zeros(ns)
    if counts:
        np.add.at(S, l, a.ravel())
        np.add.at(W, l, 1.0)
    else:
        c1 = (1 - (l - p) ** 2) ** 2
        c2 = (1 - (l + 1 - p) ** 2) ** 2
        np.add.at(S, l, c1 * a.ravel())
        np.add.at(W, l, c1)
        ok = l + 1 < ns
        np.add.at(S, l[ok] + 1, c2[ok] * a.ravel()[ok])
        np.add.at(W, l[ok] + 1, c2[ok])
    m = W != 0
    S[m] /= W[m]
    if n6:
        S /= float(nx) ** 6
    return S, W


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 16, 16), (12, 8, 10), (6, 10, 8)])
def test_spectrum_and_crossspec_brute_force(shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(sum(shape))
    f1 = rng.standard_normal(shape)
    f2 = rng.standard_normal(shape)
    S = O.spectrum_3d(O.r2c_3d(f1), nx)
    ref, W = brute_spectrum(f1)
    assert np.allclose(S, ref.real, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    sre, sim = O.crossspec_3d(O.r2c_3d(f1), O.r2c_3d(f2), nx)
    refc, _ = brute_spectrum(f1, f2)
    tol = 1e-12 * np.abs(refc).max()
    assert np.allclose(sre, refc.real, atol=tol)
    # The reference re-reads Fk(nx+2-i, j, k) for i > nx/2+1 without conjugation
    # (analysis_spec.f90:163-166), so Im S is the half-space (kx >= 0) sum with
    # multiplicity 2 for 0 < kx < nx/2 -- not the Hermitian-completed sum,
    # which vanishes by symmetry (refc.imag ~ 0 above).
    assert np.abs(refc.imag).max() < tol
    assert np.allclose(sim, brute_half_imag(f1, f2), atol=tol)


def brute_half_imag(f1, f2):
    nz, ny, nx = f1.shape
    a = np.fft.fftn(f1) * np.conj(np.fft.fftn(f2))
    a = a[:, :, : nx // 2 + 1]
    mult = np.full(nx // 2 + 1, 2.0)
    mult[0] = 1.0
    if nx % 2 == 0:
        mult[nx // 2] = 1.0
    ii = np.arange(nx // 2 + 1)
    jj = np.minimum(np.arange(ny), ny - np.arange(ny))
    kk = np.minimum(np.arange(nz), nz - np.arange(nz))
    p = np.sqrt(kk[:, None, None] ** 2 + jj[None, :, None] ** 2 + ii[None, None, :] ** 2)
    ns = int(np.floor(np.sqrt((nx // 2 + 1) ** 2 + (ny // 2 + 1) ** 2 + (nz // 2 + 1) ** 2 + 1)))
    l = np.floor(p).astype(int)
    c1 = (1 - (l - p) ** 2) ** 2 * mult
    c2 = (1 - (l + 1 - p) ** 2) ** 2 * mult
    S = np.zeros(ns)
    W = np.zeros(ns)
    np.add.at(S, l.ravel(), (c1 * a.imag).ravel())
    np.add.at(W, l.ravel(), c1.ravel())
    ok = (l + 1 < ns).ravel()
    np.add.at(S, l.ravel()[ok] + 1, (c2 * a.imag).ravel()[ok])
    np.add.at(W, l.ravel()[ok] + 1, c2.ravel()[ok])
    m = W != 0
    S[m] /= W[m]
    return S


def test_spectrum_weight_sum_rule_and_single_mode():
    n = 16
    ref_S, W = brute_spectrum(np.ones((n, n, n)))
    # sum of all two-bin weights == sum over modes of (c1 + c2), bin 0 only k = 0
    assert W[0] == 1.0
    f = plane(n, 1.0, (3, 4, 0), phase=0.0)   # |k| = 5: power only in bins 5 (c1=1)
    S = O.spectrum_3d(O.r2c_3d(f), n)
    nonzero = np.nonzero(np.abs(S) > 1e-12 * np.abs(S).max())[0]
    assert set(nonzero) == {5}


def test_discrete_spectra_brute_force():
    n = 8
    rng = np.random.default_rng(9)
    f = rng.standard_normal((n, n, n))
    F = O.r2c_3d(f)
    ref, _ = brute_spectrum(f, n6=True)
    S = O.spectrum_discrete(F, n, 0)
    m = min(len(S), len(ref))
    assert np.allclose(S[:m], ref.real[:m], rtol=1e-12, atol=1e-14)
    refc, _ = brute_spectrum(f, counts=True)
    S1 = O.spectrum_discrete(F, n, 1)
    assert np.allclose(S1, refc.real, rtol=1e-12)


def brute_half(F1, F2, n, counts=False, skip_nyq=False, n6=False):
    """Binning over the stored half space of numpy rfftn arrays (no mirror)."""
    a = (F1 * np.conj(F2)).copy()
    ii = np.arange(n // 2 + 1)
    jj = np.minimum(np.arange(n), n - np.arange(n))
    p = np.sqrt(jj[:, None, None] ** 2 + jj[None, :, None] ** 2 + ii[None, None, :] ** 2)
    keep = np.ones(p.shape, bool)
    if skip_nyq:
        keep[..., n // 2] = False
    p, a = p[keep], a[keep]
    ns = int(1.7320508075688772935274463415059 * (n // 2) + 2.0)
    l = np.floor(p).astype(int)
    S = np.zeros(ns, complex)
    W = np.zeros(ns)
    if counts:
        np.add.at(S, l, a)
        np.add.at(W, l, 1.0)
    else:
        c1, c2 = (1 - (l - p) ** 2) ** 2, (1 - (l + 1 - p) ** 2) ** 2
        np.add.at(S, l, c1 * a)
        np.add.at(W, l, c1)
        ok = l + 1 < ns
        np.add.at(S, l[ok] + 1, c2[ok] * a[ok])
        np.add.at(W, l[ok] + 1, c2[ok])
    m = W != 0
    S[m] /= W[m]
    return S / float(n) ** 6 if n6 else S


def test_discrete_nofft_spectra_brute_force():
    """spectrum_nofft (weights, /n^6, all stored x), spectrum_bin_nofft (counts,
    x-Nyquist skipped), crossspec_nofft (weights, x-Nyquist skipped, no /n^6)."""
    n = 8
    rng = np.random.default_rng(19)
    F1 = np.fft.rfftn(rng.standard_normal((n, n, n)))
    F2 = np.fft.rfftn(rng.standard_normal((n, n, n)))
    S0 = O.spectrum_discrete_k(F1, n, 0)
    assert np.allclose(S0, brute_half(F1, F1, n, n6=True).real, rtol=1e-12, atol=0)
    S1 = O.spectrum_discrete_k(F1, n, 1)
    assert np.allclose(S1, brute_half(F1, F1, n, counts=True, skip_nyq=True).real, rtol=1e-12, atol=0)
    c1, c2 = O.crossspec_discrete_k(F1, F2, n)
    ref = brute_half(F1, F2, n, skip_nyq=True)
    assert np.allclose(c1, ref.real, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    assert np.allclose(c2, ref.imag, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # the x-Nyquist plane enters spectrum_nofft only
    G = F1.copy()
    G[..., n // 2] *= 3.0
    assert not np.allclose(O.spectrum_discrete_k(G, n, 0), S0)
    assert np.array_equal(O.spectrum_discrete_k(G, n, 1), S1)


# ---------------------------------------------------------------- evolution

def hom_sim(time_var, model, mp, nfld, phi0, pi0, n=4, L=5.0):
    sim = O.Simulation(n, nfld, L, time_var, model=model, mp=mp)
    for l in range(nfld):
        sim.fld[l] = phi0[l]
        sim.fldp[l] = pi0[l]
    sim.yscl = 1.0
    sim.calc_metric()
    return sim


def hom_rhs(time_var, model, mp):
    """Hamilton's equations of the homogeneous system, from the Hamiltonians
    H = -p^2/12 + pi^2/(2a^2) + a^4 V (conformal, y = a) and
    H = -3 p^2/16 + pi^2/(2y^2) + y^2 V (cosmic, y = a^{3/2})."""
    m2chi = mp[2] if len(mp) > 2 else 0.0

    def V(f):
        if model == O.MODEL_PHI4:
            return 0.25 * mp[0] * f[0] ** 4
        return 0.5 * mp[0] * f[0] ** 2 + 0.5 * mp[1] * f[0] ** 2 * f[1] ** 2 + 0.5 * m2chi * f[1] ** 2

    def dV(f):
        if model == O.MODEL_PHI4:
            return np.array([mp[0] * f[0] ** 3])
        return np.array([mp[0] * f[0] + mp[1] * f[0] * f[1] ** 2, mp[1] * f[0] ** 2 * f[1] + m2chi * f[1]])

    def rhs(t, u):
        nf = (len(u) - 2) // 2
        f, p, y, py = u[:nf], u[nf:2 * nf], u[-2], u[-1]
        if time_var == O.CONFORMAL:
            return np.concatenate([p / y**2, -y**4 * dV(f), [-py / 6.0],
                                   [np.sum(p**2) / y**3 - 4 * y**3 * V(f)]])
        return np.concatenate([p / y**2, -y**2 * dV(f), [-3.0 * py / 8.0],
                               [np.sum(p**2) / y**3 - 2 * y * V(f)]])
    return rhs


CASES = [(O.CONFORMAL, O.MODEL_PHI4, (1.0,), 1, [1.0], [-1.1547]),
         (O.COSMIC, O.MODEL_TWOFIELD, (1.0, 4.0), 2, [1.0093430384226379, 0.3], [-0.7137133070120812, 0.1])]
# the Discrete tree's chi mass (modelv = (m2phi + g2 psi^2) phi^2 + m2psi psi^2,
# Discrete/Entropy/model.f90:35-55)
CASE_MCHI = (O.COSMIC, O.MODEL_TWOFIELD, (1.0, 4.0, 2.5), 2, [1.0093430384226379, 0.3], [-0.7137133070120812, 0.1])


@pytest.mark.parametrize("case", CASES + [CASE_MCHI], ids=["conformal_phi4", "cosmic_twofield", "cosmic_twofield_mchi"])
def test_homogeneous_limit_matches_hamilton_ode(case):
    tv, model, mp, nf, f0, p0 = case
    sim = hom_sim(tv, model, mp, nf, f0, p0)
    u0 = np.concatenate([f0, p0, [sim.yscl, sim.ysclp]])
    T, dt = 1.0, 1e-3
    sim.symp(4, dt, int(round(T / dt)))
    sol = solve_ivp(hom_rhs(tv, model, mp), (0, T), u0, rtol=1e-13, atol=1e-14, method="DOP853")
    ue = sol.y[:, -1]
    got = np.concatenate([[sim.fld[l].flat[0] for l in range(nf)], [sim.fldp[l].flat[0] for l in range(nf)],
                          [sim.yscl, sim.ysclp]])
    assert np.allclose(got, ue, rtol=1e-9, atol=1e-10), (got, ue)


def _order_errors(run, dts, T=1.0):
    tv, model, mp, nf, f0, p0 = CASES[1]
    errs = []
    for dt in dts:
        sim = hom_sim(tv, model, mp, nf, f0, p0)
        u0 = np.concatenate([f0, p0, [sim.yscl, sim.ysclp]])
        run(sim, dt, int(round(T / dt)))
        got = np.array([sim.fld[0].flat[0], sim.fldp[0].flat[0], sim.yscl])
        sol = solve_ivp(hom_rhs(tv, model, mp), (0, T), u0, rtol=1e-13, atol=1e-15, method="DOP853")
        errs.append(np.max(np.abs(got - sol.y[[0, nf, 2 * nf], -1])))
    return errs


# symp8's asymptotic order is 4: the reference (integrator.f90:73-110,
# evolve_spectral.f90:129-167) applies the 15-stage coefficients with w7
# outermost, the reverse of the stage order they were derived for.
@pytest.mark.parametrize("order,expect", [(2, 2), (4, 4), (6, 6), (8, 4)])
def test_convergence_order(order, expect):
    dts = [0.1, 0.05] if order < 8 else [0.025, 0.0125]
    errs = _order_errors(lambda s, dt, n: s.symp(order, dt, n), dts)
    ratio = errs[0] / errs[1]
    assert 0.6 * 2**expect < ratio < 1.7 * 2**expect, (order, errs, ratio)


def test_symp8_coefficients_are_eighth_order_in_forward_stage_order():
    """Same coefficients, stage order w1 (outermost) ... w7, w0, w7 ... w1,
    composed from the oracle's own symp_o2step: 8th order."""
    w = [0.74167036435061295344822780, -0.40910082580003159399730010, 0.19075471029623837995387626,
         -0.57386247111608226665638733, 0.29906418130365592384446354, 0.33462491824529818378495798,
         0.31529309239676659663205666]
    w0 = 1.0 - 2.0 * sum(w)
    seq = w + [w0] + w[::-1]

    def run(sim, dt, n):
        sim.split(seq[0] * dt / 2.0, 1)
        for j in range(n):
            for q in range(14):
                sim.symp_o2step(dt, seq[q], seq[q + 1])
            sim.symp_o2step(dt, seq[14], 0.0 if j == n - 1 else seq[0])

    errs = _order_errors(run, [0.2, 0.1])
    assert errs[0] / errs[1] > 0.6 * 2**8, errs


def test_time_reversibility():
    n, L = 8, 6.0
    rng = np.random.default_rng(3)
    sim = O.Simulation(n, 2, L, O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, 10.0))
    sim.fld[0] = 1.0 + 1e-3 * rng.standard_normal((n, n, n))
    sim.fld[1] = 1e-3 * rng.standard_normal((n, n, n))
    sim.fldp[0] = -0.7
    sim.yscl = 1.0
    sim.calc_metric()
    f0, p0, y0, py0 = sim.fld.copy(), sim.fldp.copy(), sim.yscl, sim.ysclp
    sim.symp(4, 0.01, 20)
    assert np.abs(sim.fld - f0).max() > 1e-3
    sim.symp(4, -0.01, 20)
    assert np.abs(sim.fld - f0).max() < 1e-12 and np.abs(sim.fldp - p0).max() < 1e-12
    assert abs(sim.yscl - y0) < 1e-13 and abs(sim.ysclp - py0) < 1e-12


@pytest.mark.parametrize("lap", [O.LAP_SPECTRAL, O.LAP_STENCIL])
def test_free_field_modes_rotate(lap):
    """Massive free field (g^2 = 0) with negligible energy density: H ~ 0, every
    Fourier mode rotates with omega^2 = m^2 + k_eff^2 (k_eff from the stencil
    eigenvalue for the FD Laplacian)."""
    n, L, m2 = 8, 8.0, 1.0
    rng = np.random.default_rng(11)
    stencil = (-6.0, 1.0, 0.0, 0.0, 1.0)
    sim = O.Simulation(n, 2, L, O.COSMIC, lap_kind=lap, stencil=stencil, model=O.MODEL_TWOFIELD, mp=(m2, 0.0))
    f = 1e-7 * rng.standard_normal((n, n, n))
    p = 1e-7 * rng.standard_normal((n, n, n))
    sim.fld[0], sim.fldp[0] = f, p
    sim.yscl = 1.0
    sim.calc_metric()
    T, dt = 2.0, 2e-3
    sim.symp(4, dt, int(round(T / dt)))
    dk, dx = 2 * np.pi / L, L / n
    ii = np.fft.fftfreq(n, 1.0 / n)
    if lap == O.LAP_SPECTRAL:
        kx = np.abs(ii).copy()
        kx[n // 2] = n // 2
        k2 = dk**2 * (kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kx[None, None, :] ** 2)
    else:
        s = np.sin(np.pi * ii / n) ** 2
        k2 = 4.0 / dx**2 * (s[:, None, None] + s[None, :, None] + s[None, None, :])
    w = np.sqrt(m2 + k2)
    F, Pk = np.fft.fftn(f), np.fft.fftn(p)
    Fe = F * np.cos(w * T) + Pk * np.sin(w * T) / w
    fe = np.fft.ifftn(Fe).real
    assert np.linalg.norm(sim.fld[0] - fe) < 2e-5 * np.linalg.norm(fe)


def test_friedmann_constraint_conserved():
    n, L = 8, 10.0
    rng = np.random.default_rng(12)
    sim = O.Simulation(n, 2, L, O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, 100.0))
    sim.fld[0] = 1.0 + 1e-4 * rng.standard_normal((n, n, n))
    sim.fld[1] = 1e-4 * rng.standard_normal((n, n, n))
    sim.fldp[0] = -0.7
    sim.yscl = 1.0
    sim.calc_metric()
    assert abs(sim.energies()["resid"]) < 1e-14
    sim.symp(4, 2e-3, 500)
    e = sim.energies()
    assert abs(e["resid"]) < 1e-8
    assert e["a"] > 1.0  # expanding


def test_energy_components_independent_sums():
    n, L = 8, 6.0
    rng = np.random.default_rng(13)
    sim = O.Simulation(n, 2, L, O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, 7.0))
    sim.fld[:] = rng.standard_normal(sim.fld.shape)
    sim.fldp[:] = rng.standard_normal(sim.fldp.shape)
    sim.yscl, sim.ysclp = 1.3, -0.4
    e = sim.energies()
    a = 1.3 ** (2.0 / 3.0)
    ke = 0.5 * np.mean(np.sum(sim.fldp**2, axis=0)) / 1.3**4
    pe = np.mean(0.5 * sim.fld[0] ** 2 + 0.5 * 7.0 * sim.fld[0] ** 2 * sim.fld[1] ** 2)
    dk = 2 * np.pi / L
    ii = np.abs(np.fft.fftfreq(n, 1.0 / n))
    ii[n // 2] = n // 2
    k2 = dk**2 * (ii[:, None, None] ** 2 + ii[None, :, None] ** 2 + ii[None, None, :] ** 2)
    ge = sum(0.5 * np.sum(k2 * np.abs(np.fft.fftn(sim.fld[l])) ** 2) / n**6 for l in range(2)) / a**2
    assert abs(e["ke"] - ke) < 1e-13 * ke and abs(e["pe"] - pe) < 1e-13 * pe
    assert abs(e["ge"] - ge) < 1e-12 * ge
    assert abs(e["hubble"] - (0.4 / 1.3 / 4.0)) < 1e-15


@pytest.mark.parametrize("lap", [O.LAP_SPECTRAL, O.LAP_STENCIL], ids=["spectral", "fd7"])
def test_conformal_free_massless_modes_static_background(lap):
    """Conformal time with V = 0 (lambda = 0) and a field amplitude of 1e-7,
    so the background stays at a = a0 = 1.7 (its momentum moves by O(1e-14)):
    the kick pi += h a^2 lap phi and the drift phi += h pi / a^2 give
    phi'' = lap phi, every mode rotating with omega = |k| (spectral) or the
    7-point eigenvalue (FD), independent of a0.  A wrong a-power in either
    (hamiltonian_conformal.f90:98,183-186) would give omega = a0^m |k|."""
    n, L, a0 = 8, 8.0, 1.7
    rng = np.random.default_rng(31)
    stencil = (-6.0, 1.0, 0.0, 0.0, 1.0)
    sim = O.Simulation(n, 1, L, O.CONFORMAL, lap_kind=lap, stencil=stencil, model=O.MODEL_PHI4, mp=(0.0,))
    f = 1e-7 * rng.standard_normal((n, n, n))
    p = 1e-7 * rng.standard_normal((n, n, n))
    sim.fld[0], sim.fldp[0] = f, p
    sim.yscl, sim.ysclp = a0, 0.0
    T, dt = 2.0, 2e-3
    sim.symp(4, dt, int(round(T / dt)))
    assert abs(sim.yscl - a0) < 1e-12
    dk, dx = 2 * np.pi / L, L / n
    ii = np.fft.fftfreq(n, 1.0 / n)
    if lap == O.LAP_SPECTRAL:
        kx = np.abs(ii).copy()
        kx[n // 2] = n // 2
        k2 = dk**2 * (kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kx[None, None, :] ** 2)
    else:
        s = np.sin(np.pi * ii / n) ** 2
        k2 = 4.0 / dx**2 * (s[:, None, None] + s[None, :, None] + s[None, None, :])
    w = np.sqrt(k2)
    sinc = np.where(w > 0, np.sin(w * T) / np.where(w > 0, w, 1.0), T)
    F, Pk = np.fft.fftn(f), np.fft.fftn(p) / a0**2
    fe = np.fft.ifftn(F * np.cos(w * T) + Pk * sinc).real
    assert np.linalg.norm(sim.fld[0] - fe) < 2e-5 * np.linalg.norm(fe)
    # the same run with omega scaled by a0 (a wrong a-power) is far off
    fw = np.fft.ifftn(F * np.cos(a0 * w * T) + Pk * np.where(w > 0, np.sin(a0 * w * T) / np.where(w > 0, a0 * w, 1.0),
                                                               T)).real
    assert np.linalg.norm(sim.fld[0] - fw) > 1e-1 * np.linalg.norm(fe)


def test_conformal_friedmann_constraint_with_gradient_energy():
    """Conformal lambda phi^4 with strong fluctuations (gradient energy a
    sizeable part of rho) while a grows from 1: the Friedmann residual
    (rho - 3H^2)/rho starts at 0 (calc_metric) and its violation converges
    at 4th order in dt.  The a-momentum update uses GE2 / a^2 and 4 PE
    (hamiltonian_conformal.f90:203-205) and GE in the energies carries 1/a^2
    (evolve_spectral.f90:402): a wrong power of a in either leaves an O(1)
    violation that does not shrink with dt."""
    n, L = 8, 5.0
    rng = np.random.default_rng(41)
    f0 = 1.0 + 0.05 * rng.standard_normal((n, n, n))
    p0 = -1.15 + 0.05 * rng.standard_normal((n, n, n))
    res = []
    for dt in (0.02, 0.01):
        sim = O.Simulation(n, 1, L, O.CONFORMAL, model=O.MODEL_PHI4, mp=(1.0,))
        sim.fld[0], sim.fldp[0] = f0, p0
        sim.yscl = 1.0
        sim.calc_metric()
        e0 = sim.energies()
        assert abs(e0["resid"]) < 1e-14
        assert e0["ge"] > 0.02 * e0["rho"]
        sim.symp(4, dt, int(round(1.0 / dt)))
        e = sim.energies()
        assert e["a"] > 1.3
        res.append(abs(e["resid"]))
    assert res[0] < 1e-6
    assert 10.0 < res[0] / res[1] < 24.0, res


def test_conformal_energy_components_independent_sums():
    """dump_rho in conformal time (evolve_spectral.f90:305-415 with
    hamiltonian_conformal.f90:242-261): KE = <pi^2>/2 / a^6, PE = <V>,
    GE = sum k^2 |phi_k|^2 / (2 N^6) / a^2 (Parseval), H = -p_a / (6 a^2),
    at a = 1.6."""
    n, L, lam, a, pa = 8, 6.0, 0.7, 1.6, -0.9
    rng = np.random.default_rng(43)
    sim = O.Simulation(n, 1, L, O.CONFORMAL, model=O.MODEL_PHI4, mp=(lam,))
    sim.fld[:] = rng.standard_normal(sim.fld.shape)
    sim.fldp[:] = rng.standard_normal(sim.fldp.shape)
    sim.yscl, sim.ysclp = a, pa
    e = sim.energies()
    ke = 0.5 * np.mean(sim.fldp[0] ** 2) / a**6
    pe = np.mean(0.25 * lam * sim.fld[0] ** 4)
    dk = 2 * np.pi / L
    ii = np.abs(np.fft.fftfreq(n, 1.0 / n))
    ii[n // 2] = n // 2
    k2 = dk**2 * (ii[:, None, None] ** 2 + ii[None, :, None] ** 2 + ii[None, None, :] ** 2)
    ge = 0.5 * np.sum(k2 * np.abs(np.fft.fftn(sim.fld[0])) ** 2) / n**6 / a**2
    hub = -pa / 6.0 / a**2
    assert abs(e["ke"] - ke) < 1e-13 * ke and abs(e["pe"] - pe) < 1e-13 * pe
    assert abs(e["ge"] - ge) < 1e-12 * ge
    assert abs(e["hubble"] - hub) < 1e-15
    assert abs(e["rho"] - (ke + pe + ge)) < 1e-13 * e["rho"]
    assert abs(e["resid"] - (ke + pe + ge - 3 * hub**2) / (ke + pe + ge)) < 1e-12


# ---------------------------------------------------------------- IC sampler

@pytest.mark.parametrize("n", [1, 3, 8, 17, 64])
def test_dst2_matches_direct_sum_and_scipy(n):
    """FFTW RODFT10, the sampler's radial transform (evolve_spectral.f90:457-459):
    Y_k = 2 sum_j X_j sin(pi (j + 1/2)(k + 1) / n), against the O(n^2) sum and
    scipy.fft.dst(type=2) (the same unnormalised definition)."""
    from scipy.fft import dst

    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    j = np.arange(n)
    direct = 2.0 * np.sin(np.pi * np.outer(j + 1, j + 0.5) / n) @ x
    got = O.dst2(x)
    assert np.abs(got - direct).max() <= 1e-12 * max(1.0, np.abs(direct).max())
    assert np.abs(got - dst(x, type=2)).max() <= 1e-12 * max(1.0, np.abs(direct).max())


VAC_CASES = [(16, 10.0, -0.25, 1.0, 2e5), (32, 20.0, -0.25, 3.0, 1e7 / 3.0), (16, 10.0, 0.25, 1.0, 2e5),
             (24, 7.0, -0.25, 50.0, 1.0)]


def _vacuum_profile(n, L, gamma, m2, mpl):
    """Continuum transform of the sampler's kernel: the inverse 3-D Fourier
    transform of P(k) = (k^2 + m2)^gamma exp(-(k/kcut)^2), kcut = nn dk / 2,
    radial by the DST, scaled by norm, sampled on the lattice and
    transformed back is P(k) dk^(3/2) / (4 pi sqrt(2 pi) mpl), times the
    phase exp(-i k.x0) of the kernel's centre x0 = (n/2) dx per axis."""
    dk = 2 * np.pi / L
    kz = np.fft.fftfreq(n, 1.0 / n)
    kx = np.arange(n // 2 + 1)
    k = dk * np.sqrt(kz[:, None, None] ** 2 + kz[None, :, None] ** 2 + kx[None, None, :] ** 2)
    kcut = (n // 2 + 1) * dk / 2
    P = (k**2 + m2) ** gamma * np.exp(-(k / kcut) ** 2)
    sign = (-1.0) ** (np.abs(kz[:, None, None]) + np.abs(kz[None, :, None]) + kx[None, None, :])
    return k, kcut, sign * P * dk**1.5 / (4 * np.pi * np.sqrt(2 * np.pi) * mpl)


@pytest.mark.parametrize("case", VAC_CASES, ids=["16", "32-cfg0", "16-momentum", "24-heavy"])
def test_sample_kernel_matches_continuum_profile(case):
    """The kernel's lattice transform against the continuum transform of
    the radial profile for every mode up to the cutoff: pins the DST, the
    norm constant (evolve_spectral.f90:447-448), the exponent gamma, m2eff,
    the cutoff, the division by k and the linear interpolation with the
    centre extrapolation.  A wrong factor in norm, or gamma of the wrong
    sign, is off by far more than the 0.5 % bound."""
    n, L, gamma, m2, mpl = case
    G = np.fft.rfftn(O.sample_kernel(n, L, gamma, m2, mpl))
    k, kcut, pred = _vacuum_profile(n, L, gamma, m2, mpl)
    m = k <= kcut
    assert np.abs(G.imag[m]).max() < 1e-6 * np.abs(pred[m]).max()  # real: centred kernel
    ratio = G.real[m] / pred[m]
    assert np.all(np.abs(ratio - 1.0) < 5e-3), (ratio.min(), ratio.max())


def test_sample_power_matches_vacuum_statistics():
    """Sampled fields over 40 seeds: E |R2C(f)_k|^2 = N^6 * 2 |G_k|^2 for
    modes with 0 < kx < N/2 (Rayleigh amplitude sqrt(-2 ln u): E amp^2 = 2,
    uniform phase) and N^6 |G_k|^2 on the kx = 0 and kx = N/2 planes, whose
    non-Hermitian noise the C2R symmetrises (the imaginary parts of the x-DC
    and x-Nyquist coefficients are dropped after the y/z passes).  So the
    sampled vacuum has E |phi_k|^2 = N^6 dk^3 / ((2 pi)^3 2 omega_k mpl^2)
    (gamma = -1/4) in the interior.  Bounds: four standard deviations of the
    seed average plus the kernel's 0.5 %."""
    n, L, gamma, m2, mpl = 16, 10.0, -0.25, 1.0, 2e5
    k, kcut, pred = _vacuum_profile(n, L, gamma, m2, mpl)
    acc = np.zeros(k.shape)
    M = 40
    for seed in range(M):
        acc += np.abs(np.fft.rfftn(O.sample(n, L, gamma, m2, mpl, 900 + seed, 0))) ** 2
    ratio = acc / M / (n**6 * np.abs(pred) ** 2)
    inner = (k <= kcut) & (k > 0)
    interior = inner.copy()
    interior[:, :, 0] = False
    interior[:, :, n // 2] = False
    planes = inner & ~interior
    r_int, r_pl = ratio[interior].mean(), ratio[planes].mean()
    assert abs(r_int - 2.0) < 2.0 * (4.0 / np.sqrt(M * interior.sum()) + 5e-3), r_int
    # pairs k, -k of a plane share one symmetrised value: half as many samples
    assert abs(r_pl - 1.0) < 4.0 * np.sqrt(2.0 / (M * planes.sum())) + 5e-3 + 0.02, r_pl
    # gamma = -1/4: the interior power is the vacuum 1/(2 omega)
    kk = k[interior]
    vac = n**6 * (2 * np.pi / L) ** 3 / ((2 * np.pi) ** 3 * 2.0 * np.sqrt(kk**2 + m2) * mpl**2)
    vac *= np.exp(-2.0 * (kk / kcut) ** 2)
    got = acc[interior] / M
    assert abs(np.mean(got / vac) - 1.0) < 4.0 / np.sqrt(M * interior.sum()) + 5e-3


def test_sample_statistics_and_determinism():
    n, L = 16, 10.0
    a = O.sample(n, L, -0.25, 1.0, 1.0, 5, 0)
    b = O.sample(n, L, -0.25, 1.0, 1.0, 5, 0)
    c = O.sample(n, L, -0.25, 1.0, 1.0, 5, 1)
    assert np.array_equal(a, b) and not np.allclose(a, c)
    u1, u2 = O.philox_uniforms(5, 0, 20000)
    assert np.all((u1 > 0) & (u1 <= 1)) and np.all((u2 >= 0) & (u2 < 1))
    assert abs(u1.mean() - 0.5) < 0.01 and abs(u2.var() - 1 / 12) < 0.003


def test_philox_known_answer():
    """Philox-4x32-10 known-answer vector (Salmon et al. 2011, Random123 kat_vectors):
    counter 0, key 0 -> 6627e8d5 e169c58d bc57ac4c 9b00dbd8."""
    u1, u2 = O.philox_uniforms(0, 0, 1)
    w = [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    a = ((w[0] >> 5) << 26) | (w[1] >> 6)
    b = ((w[2] >> 5) << 26) | (w[3] >> 6)
    assert u1[0] == 1.0 - a / 2.0**53 and u2[0] == b / 2.0**53


# ---------------------------------------------------------------- GW (textbook pins)

def test_tt_projector_invariants():
    n = 8
    rng = np.random.default_rng(21)
    U = rng.standard_normal((6, n, n, n // 2 + 1)) + 1j * rng.standard_normal((6, n, n, n // 2 + 1))
    P = O.tt_project(U, n)
    P2 = O.tt_project(P, n)
    assert np.abs(P2 - P).max() < 1e-12 * np.abs(P).max()  # idempotent
    idx = np.arange(n)
    ks = np.where(idx <= n // 2, idx, idx - n)
    kz, ky, kx = np.meshgrid(ks, ks, np.arange(n // 2 + 1), indexing="ij")
    kv = [kx, ky, kz]
    comp = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    M = lambda i, j: P[comp[(min(i, j), max(i, j))]]
    for i in range(3):  # transverse
        t = sum(kv[j] * M(i, j) for j in range(3))
        assert np.abs(t).max() < 1e-12 * np.abs(P).max()
    assert np.abs(M(0, 0) + M(1, 1) + M(2, 2)).max() < 1e-12 * np.abs(P).max()  # traceless


def test_gw_source_plane_wave():
    n, L = 16, 6.0
    dk = 2 * np.pi / L
    kv = (1, 2, -1)
    f = plane(n, L, kv)
    S = O.gw_source(np.stack([f, np.zeros_like(f)]), dk)
    x = np.arange(n) * L / n
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    s2 = np.sin(dk * (kv[0] * x3 + kv[1] * y3 + kv[2] * z3) + 0.3) ** 2
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for c, (i, j) in enumerate(pairs):
        assert np.abs(S[c] - dk**2 * kv[i] * kv[j] * s2).max() < 1e-11


def test_gw_source_fd_plane_wave():
    """Central differences of cos(k.x + 0.3): d_a f = -sin(k.x + 0.3) sin(k_a dx) / dx exactly."""
    n, L = 16, 6.0
    dk, dx = 2 * np.pi / L, L / n
    kv = (1, 2, -1)
    f = plane(n, L, kv)
    S = O.gw_source_fd(np.stack([f, 2.0 * f]), dx)
    x = np.arange(n) * L / n
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    s2 = np.sin(dk * (kv[0] * x3 + kv[1] * y3 + kv[2] * z3) + 0.3) ** 2
    g = [np.sin(kv[a] * dk * dx) / dx for a in range(3)]
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for c, (i, j) in enumerate(pairs):
        assert np.abs(S[c] - 5.0 * g[i] * g[j] * s2).max() < 1e-11  # 1 + 2^2 from the two fields


def _gw_sim(n, L, lap, time_var=O.CONFORMAL, nfld=1, model=O.MODEL_PHI4, mp=(0.0,), **kw):
    return O.Simulation(n, nfld, L, time_var, lap_kind=lap, model=model, mp=mp, gw=True, **kw)


@pytest.mark.parametrize("lap", [O.LAP_SPECTRAL, O.LAP_STENCIL], ids=["spectral", "fd7"])
def test_gw_free_wave_static_background(lap):
    """No fields, a = 1, p_a = 0: the background stays static and each u_ij
    mode is a free wave, u(t) = A cos(w t) cos(k.x) with w^2 = k^2 (spectral)
    or the 7-point eigenvalue (FD).  Yoshida-4 error shrinks as dt^4."""
    n, L = 16, 8.0
    dk, dx = 2 * np.pi / L, L / n
    kv = np.array([2, 1, -1])
    x = np.arange(n) * dx
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    mode = np.cos(dk * (kv[0] * x3 + kv[1] * y3 + kv[2] * z3))
    if lap == O.LAP_SPECTRAL:
        w = dk * np.sqrt((kv ** 2).sum())
    else:
        w = np.sqrt((4.0 / dx ** 2) * (np.sin(kv * dk * dx / 2) ** 2).sum())
    T = 1.0
    errs = []
    for nsteps in (20, 40):
        s = _gw_sim(n, L, lap)
        s.gwu[1] = 0.5 * mode
        s.gwu[4] = -0.25 * mode
        s.symp(4, T / nsteps, nsteps)
        assert s.yscl == 1.0 and s.ysclp == 0.0
        exact = np.cos(w * T) * mode
        errs.append(max(np.abs(s.gwu[1] - 0.5 * exact).max(), np.abs(s.gwu[4] + 0.25 * exact).max()))
        assert np.abs(s.gwu[[0, 2, 3, 5]]).max() == 0.0
        assert np.abs(s.gwp[1] + 0.5 * w * np.sin(w * T) * mode).max() < 1e-3
    assert errs[0] < 1e-4
    assert 12.0 < errs[0] / errs[1] < 20.0


@pytest.mark.parametrize("time_var", [O.CONFORMAL, O.COSMIC], ids=["conformal", "cosmic"])
def test_gw_single_kick_map(time_var):
    """One kick from u = pi_u = 0: pi_u = hA * (2 / M_p^2) * S_ij exactly, with
    hA the field kick's Laplacian coefficient (a^2 h conformal, y^2 h / a^2
    cosmic) and S from numpy central differences."""
    n, L, h, mpl = 8, 5.0, 0.01, 0.7
    rng = np.random.default_rng(4)
    s = _gw_sim(n, L, O.LAP_SPECTRAL, time_var=time_var, nfld=2, model=O.MODEL_TWOFIELD, mp=(1.0, 10.0),
                gw_mpl=mpl)
    s.fld[...] = 0.1 * rng.standard_normal(s.fld.shape)
    s.fldp[...] = 0.1 * rng.standard_normal(s.fld.shape)
    s.yscl = 1.3
    s.split(h, 3)
    dx = L / n
    g = [[(np.roll(f, -1, axis=2 - a) - np.roll(f, 1, axis=2 - a)) / (2 * dx) for a in range(3)] for f in s.fld]
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    a = 1.3 if time_var == O.CONFORMAL else 1.3 ** (2.0 / 3.0)
    hA = h * 1.3 ** 2 if time_var == O.CONFORMAL else h * 1.3 ** 2 / a ** 2
    for c, (i, j) in enumerate(pairs):
        Sij = sum(gl[i] * gl[j] for gl in g)
        assert np.abs(s.gwp[c] - hA * (2.0 / mpl ** 2) * Sij).max() < 1e-14 * max(1.0, np.abs(Sij).max())
    assert np.abs(s.gwu).max() == 0.0


def test_gw_one_dimensional_source_has_no_tt_part():
    """A field varying along x only sources only u_xx, whose modes all have k
    along x: the TT projection (and the GW spectrum) vanish identically."""
    n, L = 16, 6.0
    x = np.arange(n) * L / n
    s = _gw_sim(n, L, O.LAP_SPECTRAL, time_var=O.COSMIC, nfld=2, model=O.MODEL_TWOFIELD, mp=(1.0, 10.0))
    prof = 0.2 * np.cos(2 * np.pi * x / L) + 0.05 * np.sin(6 * np.pi * x / L)
    s.fld[0] = 1.0 + prof[None, None, :]
    s.fld[1] = 0.3 * prof[None, None, :]
    s.calc_metric()
    s.symp(4, 0.01, 10)
    assert np.abs(s.gwu[0]).max() > 1e-6
    assert np.abs(s.gwu[1:]).max() == 0.0
    Uk = np.stack([O.r2c_3d(s.gwp[c]) for c in range(6)])
    TT = O.tt_project(Uk, n)
    assert np.abs(TT).max() < 1e-12 * np.abs(Uk).max()


# ---------------------------------------------------------------- Discrete make_output

def _np_discrete(fld, fldp, c, dx, a, hub, mp):
    """Discrete/Entropy get_variables + output_reg + output_log_file in numpy:
    stencil sums with np.roll over the 26 neighbours (rank = number of nonzero
    offsets), per-site rho and P, their statistics, the two fields and the
    half-space spectra of their transforms."""
    nf = fld.shape[0]
    offs = [(dz, dy, dx_) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx_ in (-1, 0, 1)]
    rank = lambda o: sum(1 for v in o if v)
    sh = lambda f, o: np.roll(f, shift=(-o[0], -o[1], -o[2]), axis=(0, 1, 2))  # f(x + o)
    lap = [c[0] * fld[l] for l in range(nf)]
    g2 = np.zeros(fld.shape[1:])
    pp = [np.zeros(fld.shape[1:]) for _ in range(nf)]
    for o in offs:
        r = rank(o)
        if r == 0:
            continue
        for l in range(nf):
            fn, pn = sh(fld[l], o), sh(fldp[l], o)
            lap[l] = lap[l] + c[r] * fn
            g2 = g2 + c[r] * (fn - fld[l]) ** 2
            pp[l] = pp[l] + c[r] * (pn - fldp[l]) * (fn - fld[l])
    T = np.sum(fldp ** 2, axis=0)
    if nf == 1:
        V = 2.0 * 0.25 * mp[0] * fld[0] ** 4
    else:
        V = 2.0 * (0.5 * mp[0] * fld[0] ** 2 + 0.5 * mp[1] * fld[0] ** 2 * fld[1] ** 2 + 0.5 * mp[2] * fld[1] ** 2)
    ekin, egrad, epot = 0.5 / a ** 6, 0.25 / (a * dx) ** 2 / c[4], 0.5
    rho = ekin * T + egrad * g2 + epot * V
    prs = ekin * T - egrad * g2 / 3.0 - epot * V
    KE, GE, PE = ekin * T.mean(), egrad * g2.mean(), epot * V.mean()
    rhoave, prsave = KE + GE + PE, KE - PE - GE / 3.0
    lnrhoave = np.log(rho).mean() - np.log(rhoave)
    grav = -3.0 * hub ** 2
    log = [a, hub, KE + GE + PE, prsave, lnrhoave, (prs / rho).mean(), KE, PE, GE, grav, KE + GE + PE + grav]
    log += [fld[l].mean() for l in range(nf)] + [fldp[l].mean() / a ** 3 for l in range(nf)]
    lapc, ppc = 1.0 / (a * dx) ** 2 / c[4], 0.5 / (a * dx) ** 2 / c[4]
    num = sum(fldp[l] * lapc * lap[l] for l in range(nf)) + sum(ppc * pp[l] for l in range(nf))
    lnr = np.log(rho) - np.log(rhoave) - lnrhoave
    dln = -3.0 * hub * (prs / rho - prsave / rhoave) + num / rho / a ** 3
    n = fld.shape[1]
    ns = int(np.sqrt(3.0) * (n // 2) + 2)
    F1, F2 = np.fft.rfftn(lnr), np.fft.rfftn(dln)
    kz = np.abs(np.fft.fftfreq(n, 1.0 / n))
    kz[n // 2] = n // 2
    kx = np.arange(n // 2 + 1)
    p = np.sqrt(kz[:, None, None] ** 2 + kz[None, :, None] ** 2 + kx[None, None, :] ** 2)
    keep = kx[None, None, :] < n // 2 + 0 * p
    l = np.floor(p).astype(int)
    c1, c2 = (1 - (l - p) ** 2) ** 2, (1 - (l + 1 - p) ** 2) ** 2
    x = F2 * np.conj(F1)
    vals = [np.abs(F1) ** 2, np.abs(F2) ** 2, x.real, x.imag]
    psd = np.zeros((4, ns))
    W = np.zeros(ns)
    m = keep
    np.add.at(W, l[m], c1[m])
    m2 = m & (l + 1 < ns)
    np.add.at(W, l[m2] + 1, c2[m2])
    for ch, v in enumerate(vals):
        np.add.at(psd[ch], l[m], (c1 * v)[m])
        np.add.at(psd[ch], l[m2] + 1, (c2 * v)[m2])
    psd = np.where(W > 0, psd / np.where(W > 0, W, 1.0), psd)
    return np.array(log), lnr, dln, psd


@pytest.mark.parametrize("stencil,nf", [((-6.0, 1.0, 0.0, 0.0, 1.0), 2), ((-128.0, 14.0, 3.0, 1.0, 30.0), 2),
                                        ((-6.0, 1.0, 0.0, 0.0, 1.0), 1)], ids=["7pt", "27pt", "7pt-phi4"])
def test_discrete_output_matches_numpy(stencil, nf):
    """orc_discrete_output (get_variables, output_reg, output_log_file of
    Discrete/Entropy/iofile.f90) against a numpy restatement built on np.roll
    stencils and np.fft.rfftn, on a random state with y != 1 and p_y != 0."""
    n, L = 8, 5.0
    rng = np.random.default_rng(17 + nf)
    model = O.MODEL_TWOFIELD if nf == 2 else O.MODEL_PHI4
    mp = (1.0, 3.0, 0.5) if nf == 2 else (0.8,)
    sim = O.Simulation(n, nf, L, O.COSMIC, lap_kind=O.LAP_STENCIL, stencil=stencil, model=model, mp=mp)
    sim.fld[:] = 1.0 + 0.1 * rng.standard_normal(sim.fld.shape)
    sim.fldp[:] = 0.3 + 0.1 * rng.standard_normal(sim.fldp.shape)
    sim.yscl, sim.ysclp = 1.4, -0.6
    log, lnr, dln, psd = sim.discrete_output(2.5)
    a, hub = 1.4 ** (2.0 / 3.0), 0.6 / 1.4 / 4.0
    nlog, nlnr, ndln, npsd = _np_discrete(sim.fld, sim.fldp, stencil, L / n, a, hub, mp + (0.0,) * 3)
    assert log[0] == 2.5
    assert np.allclose(log[1:], nlog, rtol=1e-12, atol=1e-14)
    assert np.allclose(lnr, nlnr, rtol=1e-12, atol=1e-13) and np.allclose(dln, ndln, rtol=1e-11, atol=1e-12)
    assert np.allclose(psd, npsd, rtol=1e-10, atol=1e-12 * np.abs(npsd).max())
    # a homogeneous state: no gradients, uniform rho: every field statistic is
    # trivial and both transformed fields vanish
    sim.fld[:] = 1.1
    sim.fldp[:] = 0.2
    log, lnr, dln, psd = sim.discrete_output(0.0)
    assert abs(log[5]) < 1e-13 and np.abs(lnr).max() < 1e-13 and np.abs(psd).max() < 1e-20
