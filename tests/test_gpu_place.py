"""Placement tuner (lat_place_tune): the field, momentum and k-space arrays
move to the fastest of several HBM allocations with the state carried over
bit for bit, and evolution afterwards is bitwise equal to evolution without
the tuner (same kernels, launch shapes and reduction order; only the
addresses differ).  The tuned and untuned runs make the same calls in the
same order apart from the tuner itself."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PHI0 = 1.0093430384226378929425913902459
DPHI0 = -0.7137133070120812430962278466136


def _init(lat, g2=2.5e5):
    m2 = [1.0, g2 * PHI0 ** 2]
    lat.sample_field(0, 0, PHI0, -0.25, m2[0], 2e5, 20240601, 0)
    lat.sample_field(0, 1, DPHI0, 0.25, m2[0], 2e5, 20240601, 1)
    lat.sample_field(1, 0, 0.0, -0.25, m2[1], 2e5, 20240601, 2)
    lat.sample_field(1, 1, 0.0, 0.25, m2[1], 2e5, 20240601, 3)
    lat.calc_metric()


def _run(la, n, prec, lap, gw, tune, steps):
    with la.Lattice(n, nfld=2, length=10.0, precision=prec, laplacian=lap, gw=gw,
                    model_params=(1.0, 2.5e5)) as lat:
        _init(lat)
        dt = (10.0 / n) / 40.0
        lat.symp4(dt, steps)
        # spectra first: an FP64 context's transform buffer for spectra is
        # then its k-space array, which the tuner moves along with it
        pre = lat.output_spectra(0)
        before = lat.get_state()
        kept, ms = lat.place_tune(4) if tune else (None, None)
        after = lat.get_state()
        lat.symp4(dt, steps)
        out = dict(pre=pre, before=before, after=after, state=lat.get_state(), e=lat.energies(),
                   spec=[lat.output_spectra(0), lat.output_spectra(1)], gw=lat.get_gw() if gw else None)
    return out, kept, ms


def _same(x, y):
    if isinstance(x, dict):
        assert x.keys() == y.keys()
        for k in x:
            _same(x[k], y[k])
    elif isinstance(x, (list, tuple)):
        assert len(x) == len(y)
        for a, b in zip(x, y):
            _same(a, b)
    else:
        np.testing.assert_array_equal(np.asarray(x), np.asarray(y))


@pytest.mark.parametrize("precision,lap,gw", [("f64", "spectral", False), ("mixed", "spectral", False),
                                              ("f64", "fd7", False), ("mixed", "fd7", False),
                                              ("mixed", "spectral", True), ("f64", "fd7", True)],
                         ids=["f64-spectral", "mixed-spectral", "f64-fd7", "mixed-fd7", "mixed-spectral-gw",
                              "f64-fd7-gw"])
def test_place_tune_is_invisible(gpu, precision, lap, gw):
    print("case place tune 64", precision, lap, gw, flush=True)
    prec = gpu.F64 if precision == "f64" else gpu.MIXED
    lp = gpu.LAP_SPECTRAL if lap == "spectral" else gpu.LAP_STENCIL
    t, kept, ms = _run(gpu, 64, prec, lp, gw, True, 5)
    u, _, _ = _run(gpu, 64, prec, lp, gw, False, 5)
    print("kept", kept, "ms per force evaluation", ms, flush=True)
    assert 0 <= kept < 4 and len(ms) == 4 and all(x > 0 for x in ms)
    assert ms[kept] == min(ms)
    _same(t["before"], t["after"])  # the state moved unchanged
    _same(t["before"], u["before"])
    for k in ("pre", "state", "e", "spec", "gw"):
        _same(t[k], u[k])


def test_place_tune_1024_mixed(gpu):
    """At the headline size: energies, background and spectra after one step
    on each side of the tuner equal those of the untuned run bitwise."""
    print("case place tune 1024 mixed", flush=True)
    res = []
    for tune in (True, False):
        with gpu.Lattice(1024, nfld=2, length=10.0, precision=gpu.MIXED, model_params=(1.0, 2.5e5)) as lat:
            _init(lat)
            dt = (10.0 / 1024) / 40.0
            lat.symp4(dt, 1)
            e0 = lat.energies()
            kept, ms = lat.place_tune(3) if tune else (None, None)
            e1 = lat.energies()
            lat.symp4(dt, 1)
            res.append(dict(e0=e0, e1=e1, e=lat.energies(), sc=lat.scalars,
                            spec=[lat.output_spectra(0), lat.output_spectra(1)]))
            if tune:
                print("kept", kept, "ms per force evaluation", ms, flush=True)
                assert 0 <= kept < 3 and ms[kept] == min(x for x in ms if x > 0)
    t, u = res
    _same(t["e0"], t["e1"])
    _same(t, u)


def test_place_tune_rejects_batches_views_and_bad_counts(gpu):
    print("case place tune errors", flush=True)
    with gpu.Lattice(16, members=[(1.0, 2.5e5), (1.0, 1e5)]) as b:
        with pytest.raises(gpu.LatticeError):
            b.place_tune(2)
        with pytest.raises(gpu.LatticeError):
            b.member(0).place_tune(2)
    with gpu.Lattice(16) as lat:
        for k in (0, 9):
            with pytest.raises(gpu.LatticeError):
                lat.place_tune(k)
        assert lat.place_tune(1) == (0, [-1.0])
