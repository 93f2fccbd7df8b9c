"""GPU parity of the single operators (FFT, Laplacians, spectra) against the
CPU oracle and numpy.  Calls go through the C ABI (lattice_amd -> liblattice_amd.so)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# every line length of the kernels' size list along every axis: x rows (R2C /
# C2R) and y / z columns (the column passes, the fused z pass of the
# Laplacian) of 96, 128, 192 and 256, the config-1 and ensemble sizes
SIZES = [(8, 8, 8), (16, 16, 16), (32, 32, 32), (64, 64, 64), (16, 32, 8), (48, 48, 48), (96, 32, 16),
         (192, 16, 8), (128, 16, 16), (256, 8, 8), (1024, 8, 8), (8, 512, 8), (8, 1024, 8), (8, 8, 1024),
         (16, 96, 128), (8, 256, 192), (16, 128, 96), (8, 192, 256), (32, 256, 256)]


def _lat(la, shape, **kw):
    nx, ny, nz = shape
    kw.setdefault("nfld", 2)
    return la.Lattice((nx, ny, nz), length=10.0, **kw)


@pytest.mark.parametrize("shape", SIZES, ids=lambda s: "x".join(map(str, s)))
def test_r2c_matches_numpy(gpu, shape):
    print("case r2c", shape, flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 7 + ny * 3 + nz)
    f = rng.standard_normal((nz, ny, nx))
    with _lat(gpu, shape) as lat:
        F = lat.fft_r2c_3d(f)
        back = lat.fft_c2r_3d(F)
    G = np.fft.rfftn(f)
    assert np.abs(F - G).max() <= 1e-13 * np.abs(G).max() * np.log2(f.size)
    assert np.abs(back / f.size - f).max() <= 1e-13 * np.log2(f.size)


@pytest.mark.parametrize("shape", SIZES, ids=lambda s: "x".join(map(str, s)))
def test_c2r_matches_numpy_nonhermitian(gpu, shape):
    """C2R of an arbitrary half-complex array: imaginary parts of the x-DC and
    x-Nyquist coefficients are discarded after the y/z passes (numpy irfftn)."""
    print("case c2r", shape, flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(3 + nx)
    F = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
    with _lat(gpu, shape) as lat:
        f = lat.fft_c2r_3d(F)
    ref = np.fft.irfftn(F, s=(nz, ny, nx), axes=(0, 1, 2)) * (nx * ny * nz)
    assert np.abs(f - ref).max() <= 1e-13 * np.abs(ref).max() * np.log2(f.size)


@pytest.mark.parametrize("shape", SIZES, ids=lambda s: "x".join(map(str, s)))
def test_laplacian_matches_oracle(gpu, shape):
    print("case laplacian", shape, flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(11 + nx)
    f = rng.standard_normal((nz, ny, nx))
    dk = 2 * np.pi / 10.0
    with _lat(gpu, shape) as lat:
        g = lat.laplacian_3d(f, dk)
    r = O.laplacian_3d(f, dk)
    assert np.abs(g - r).max() <= 1e-13 * np.abs(r).max()


def test_laplacian_plane_wave_closed_form(gpu):
    n, L = 32, 10.0
    dk = 2 * np.pi / L
    x = np.arange(n) * L / n
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    f = np.cos(3 * dk * x3 + 5 * dk * y3 - 7 * dk * z3) + np.cos(16 * dk * x3)  # incl. x-Nyquist
    expect = -(9 + 25 + 49) * dk**2 * np.cos(3 * dk * x3 + 5 * dk * y3 - 7 * dk * z3) \
        - (16 * dk) ** 2 * np.cos(16 * dk * x3)
    with gpu.Lattice(n, nfld=2, length=L) as lat:
        g = lat.laplacian_3d(f)
    assert np.abs(g - expect).max() < 1e-11


@pytest.mark.parametrize("shape", [(8, 8, 8), (32, 32, 32), (64, 64, 64), (16, 32, 8), (48, 48, 48),
                                   (192, 16, 8)], ids=lambda s: "x".join(map(str, s)))
def test_spectra_match_oracle(gpu, shape):
    print("case spectra", shape, flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(5 + nx)
    f1 = rng.standard_normal((nz, ny, nx))
    f2 = rng.standard_normal((nz, ny, nx))
    with _lat(gpu, shape) as lat:
        S1, F1 = lat.spectrum_3d(f1, return_fk=True)
        S2, F2 = lat.spectrum_3d(f2, return_fk=True)
        cre, cim = lat.crossspec_3d(F2, F1)
        # repeatability: deterministic binning
        S1b = lat.spectrum_3d(f1)
    R1 = O.r2c_3d(f1)
    R2 = O.r2c_3d(f2)
    assert np.abs(F1 - R1).max() <= 1e-13 * np.abs(R1).max()
    o1 = O.spectrum_3d(R1, nx)
    ore, oim = O.crossspec_3d(R2, R1, nx)
    assert np.allclose(S1, o1, rtol=1e-12, atol=0)
    assert np.allclose(cre, ore, rtol=1e-10, atol=1e-12 * np.abs(ore).max())
    assert np.allclose(cim, oim, rtol=1e-10, atol=1e-12 * np.abs(ore).max())
    assert np.array_equal(S1, S1b)


@pytest.mark.parametrize("n", [8, 32, 48, 64])
@pytest.mark.parametrize("kind", [0, 1], ids=["weighted", "counts"])
def test_discrete_spectra_match_oracle(gpu, n, kind):
    """Discrete-tree spectrum (two-bin weights, /n^6) and spectrum_bin
    (integer-count bins) with ns = int(sqrt(3) n/2 + 2) bins."""
    print("case discrete spectra", n, kind, flush=True)
    rng = np.random.default_rng(11 + n + kind)
    f = rng.standard_normal((n, n, n))
    with _lat(gpu, (n, n, n), nfld=1, model=gpu.MODEL_PHI4, model_params=(1.0,)) as lat:
        S = lat.spectrum_discrete(f, kind)
        S2 = lat.spectrum_discrete(f, kind)
    ref = O.spectrum_discrete(O.r2c_3d(f), n, kind)
    assert S.shape == ref.shape == (int(np.sqrt(3.0) * (n // 2) + 2),)
    assert np.allclose(S, ref, rtol=1e-12, atol=0)
    assert np.array_equal(S, S2)


@pytest.mark.parametrize("n", [8, 32, 48])
def test_discrete_nofft_spectra_match_oracle(gpu, n):
    print("case discrete nofft spectra", n, flush=True)
    rng = np.random.default_rng(31 + n)
    F1 = np.fft.rfftn(rng.standard_normal((n, n, n)))
    F2 = np.fft.rfftn(rng.standard_normal((n, n, n)))
    with _lat(gpu, (n, n, n), nfld=1, model=gpu.MODEL_PHI4, model_params=(1.0,)) as lat:
        S0 = lat.spectrum_discrete_k(F1, 0)
        S1 = lat.spectrum_discrete_k(F1, 1)
        c1, c2 = lat.crossspec_discrete_k(F1, F2)
    assert np.allclose(S0, O.spectrum_discrete_k(F1, n, 0), rtol=1e-12, atol=0)
    assert np.allclose(S1, O.spectrum_discrete_k(F1, n, 1), rtol=1e-12, atol=0)
    r1, r2 = O.crossspec_discrete_k(F1, F2, n)
    scale = np.abs(r1).max()
    assert np.allclose(c1, r1, rtol=1e-11, atol=1e-13 * scale)
    assert np.allclose(c2, r2, rtol=1e-11, atol=1e-13 * scale)


def test_spectra_weight_cache_across_bin_sets(gpu):
    """The binning weights W(b) are binned once per (bin count, weighting
    flags) and cached in four first-in-first-out slots per lattice.  At n = 16
    the calls below carry five keys (spectrum_3d and discrete kind 0 share
    ns = 15 with flags 0): A = (15, 0), B = counts, C = half space,
    D = counts + half + no x-Nyquist, E = half + no x-Nyquist.  The order
    A B C A D B E A C D A visits: A, B, C miss (slots 0-2); A hits slot 0
    after two other slots were written; D misses (slot 3); B hits slot 1
    after two more writes; E misses and evicts A (slot 0); A misses and
    evicts B (slot 1); C hits slot 2 after three writes; D hits slot 3; A
    hits slot 1.  Every call must give bitwise what it gives on a fresh
    lattice (where its weights are binned for that call)."""
    print("case spectra weight cache", flush=True)
    n = 16
    rng = np.random.default_rng(77)
    f = rng.standard_normal((n, n, n))
    F1 = np.fft.rfftn(rng.standard_normal((n, n, n)))
    F2 = np.fft.rfftn(rng.standard_normal((n, n, n)))
    calls = {
        "A": lambda lat: lat.spectrum_3d(f),
        "A'": lambda lat: lat.spectrum_discrete(f, 0),
        "B": lambda lat: lat.spectrum_discrete(f, 1),
        "C": lambda lat: lat.spectrum_discrete_k(F1, 0),
        "D": lambda lat: lat.spectrum_discrete_k(F1, 1),
        "E": lambda lat: np.concatenate(lat.crossspec_discrete_k(F1, F2)),
    }
    fresh = {}
    for key, call in calls.items():
        with _lat(gpu, (n, n, n), nfld=1, model=gpu.MODEL_PHI4, model_params=(1.0,)) as lat:
            fresh[key] = call(lat)
    order = ["A", "B", "C", "A'", "D", "B", "E", "A", "C", "D", "A'"]
    with _lat(gpu, (n, n, n), nfld=1, model=gpu.MODEL_PHI4, model_params=(1.0,)) as lat:
        for q, key in enumerate(order):
            got = calls[key](lat)
            assert np.array_equal(got, fresh[key]), (q, key)
    # the fresh results themselves are pinned to the oracle by the tests above
    assert np.allclose(fresh["A'"], O.spectrum_discrete(O.r2c_3d(f), n, 0), rtol=1e-12, atol=0)


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_spectra_weight_cache_shared_by_batch_views(gpu, precision):
    """Member views use their batch's cache: view 0's output spectra bin the
    weights (miss), view 1's reuse them (hit); both equal single-context
    output spectra of the same state bitwise, and a later spectrum with other
    flags on view 1 does not disturb view 0's next call."""
    print("case spectra cache views", precision, flush=True)
    n, L = 16, 10.0
    prec = gpu.F64 if precision == "f64" else gpu.MIXED
    rng = np.random.default_rng(5)
    ics = []
    for m in range(2):
        fld = np.stack([1.0 + 1e-3 * rng.standard_normal((n, n, n)), 1e-3 * rng.standard_normal((n, n, n))])
        fldp = np.stack([-0.7 + 1e-3 * rng.standard_normal((n, n, n)), 1e-3 * rng.standard_normal((n, n, n))])
        ics.append((fld, fldp))
    single = []
    for m in range(2):
        with gpu.Lattice(n, nfld=2, length=L, precision=prec) as lat:
            lat.set_state(*ics[m], 1.0, 0.0)
            single.append([lat.output_spectra(0), lat.output_spectra(1)])
    with gpu.Lattice(n, nfld=2, length=L, precision=prec, members=[(1.0, 2.5e5)] * 2) as batch:
        views = [batch.member(m) for m in range(2)]
        for m, v in enumerate(views):
            v.set_state(*ics[m], 1.0, 0.0)
        got0 = [views[0].output_spectra(0), views[0].output_spectra(1)]
        got1 = [views[1].output_spectra(0), views[1].output_spectra(1)]
        views[1].spectrum_discrete(ics[1][0][0], 1)
        again0 = views[0].output_spectra(1)
        for v in views:
            v.close()
    for l in range(2):
        assert np.array_equal(got0[l], single[0][l]) and np.array_equal(got1[l], single[1][l]), l
    assert np.array_equal(again0, single[0][1])


def test_discrete_spectra_need_cubic_lattice(gpu):
    with _lat(gpu, (16, 32, 8)) as lat:
        with pytest.raises(gpu.LatticeError):
            lat.spectrum_discrete(np.zeros((8, 32, 16)), 0)


@pytest.mark.parametrize("coef", [(-6.0, 1.0, 0.0, 0.0, 1.0), (-128.0, 14.0, 3.0, 1.0, 30.0)], ids=["7pt", "27pt"])
def test_stencil_matches_oracle(gpu, coef):
    n, L = 32, 10.0
    rng = np.random.default_rng(2)
    f = rng.standard_normal((n, n, n))
    with gpu.Lattice(n, nfld=2, length=L, laplacian=gpu.LAP_STENCIL, stencil=coef) as lat:
        g = lat.stencil_laplacian(f)
    r = O.stencil_laplacian(f, coef, L / n)
    assert np.abs(g - r).max() <= 1e-12 * np.abs(r).max()


def test_gradient_squared_matches_oracle(gpu):
    n, L = 32, 10.0
    rng = np.random.default_rng(4)
    f = rng.standard_normal((n, n, n))
    with gpu.Lattice(n, nfld=2, length=L) as lat:
        g = lat.gradient_squared_3d(f)
    r = O.gradient_squared_3d(f, 2 * np.pi / L)
    assert np.abs(g - r).max() <= 1e-12 * np.abs(r).max()
