"""Separable form of the spectral Laplacian in the mixed-form kicks (x:
drift + D_xx, y: + D_yy, z: + D_zz with the kick and reductions) against the
five-sweep 3-D form on the same state, and against the FP64 form.  The
operator is the same (the 3-D multiplier is a sum of one-axis terms), so the
two forms differ by FP32 rounding only; a wrong factor, axis or pairing
gives O(1) differences."""
import numpy as np
import pytest

from test_gpu_mixed import PHI0, DPHI0, spec_err

pytestmark = pytest.mark.gpu


def sample(lat, seed, g2):
    lat.sample_field(0, 0, PHI0, -0.25, 1.0, 2e5, seed, 0)
    lat.sample_field(0, 1, DPHI0, 0.25, 1.0, 2e5, seed, 1)
    lat.sample_field(1, 0, 0.0, -0.25, g2 * PHI0**2, 2e5, seed, 2)
    lat.sample_field(1, 1, 0.0, 0.25, g2 * PHI0**2, 2e5, seed, 3)
    lat.set_scalars(1.0, 0.0)
    lat.calc_metric()


def run(gpu, prec, n, form, steps, chunks, seed=11, host_state=True):
    L, g2 = 10.0, 2.5e5
    dt = (L / n) / 40
    with gpu.Lattice(n, nfld=2, length=L, precision=prec, time_var=gpu.COSMIC, model=gpu.MODEL_TWOFIELD,
                     model_params=(1.0, g2)) as lat:
        if form:
            lat.set_laplacian_form(True)
            assert lat.laplacian_form() == 1
        sample(lat, seed, g2)
        spectra = []
        for _ in range(chunks):
            lat.symp4(dt, steps // chunks)
            spectra.append((lat.output_spectra(0), lat.output_spectra(1)))
        return lat.energies(), spectra, (lat.get_state() if host_state else None)


@pytest.mark.parametrize("n", [128, 256])
def test_separable_matches_3d_form_and_fp64(gpu, n):
    print("case separable", n, flush=True)
    steps, chunks = 24, 3
    e3, s3, st3 = run(gpu, gpu.MIXED, n, 0, steps, chunks)
    es, ss, sts = run(gpu, gpu.MIXED, n, 1, steps, chunks)
    ed, sd, std_ = run(gpu, gpu.F64, n, 0, steps, chunks)
    # background and energies: FP64 quantities fed by FP32 sweeps
    assert abs(es["a"] - e3["a"]) <= 1e-13 and abs(es["a"] - ed["a"]) <= 1e-12
    for k in ("rho", "ke", "pe", "ge"):
        assert abs(es[k] - e3[k]) <= 1e-8 * abs(e3["rho"]), k
        assert abs(es[k] - ed[k]) <= 1e-8 * abs(ed["rho"]), k
    assert abs(es["resid"]) < 1e-7
    # spectra: same bound as the 3-D mixed form against FP64
    e_sep_d, e_3d_d, e_sep_3d = spec_err(ss, sd, 1e-3), spec_err(s3, sd, 1e-3), spec_err(ss, s3, 1e-3)
    print("spectrum rel err: sep vs f64", e_sep_d, "3d vs f64", e_3d_d, "sep vs 3d", e_sep_3d, flush=True)
    assert e_sep_d < 1e-5 and e_sep_3d < 1e-5
    # fluctuation fields and momenta
    for q in (0, 1):
        for l in range(2):
            ref = std_[q][l] - std_[q][l].mean()
            assert np.linalg.norm(sts[q][l] - std_[q][l]) <= 1e-4 * np.linalg.norm(ref), (q, l)
            assert np.linalg.norm(sts[q][l] - st3[q][l]) <= 1e-4 * np.linalg.norm(ref), (q, l)


def test_separable_batch_members_match_single_runs(gpu):
    """A batch runs the same kernels with the member from the grid: each
    member equals its single-context run bitwise."""
    print("case separable batch", flush=True)
    n, L, dt = 128, 10.0, (10.0 / 128) / 40
    g2s = [1e5, 2.5e5]
    single = []
    for m, g2 in enumerate(g2s):
        with gpu.Lattice(n, nfld=2, length=L, precision=gpu.MIXED, model_params=(1.0, g2)) as lat:
            lat.set_laplacian_form(True)
            sample(lat, 20 + m, g2)
            ic = lat.get_state()
            lat.set_state(*ic)  # both runs start from the same host state
            lat.calc_metric()
            lat.symp4(dt, 3)
            single.append((ic, lat.get_state()))
    with gpu.Lattice(n, nfld=2, length=L, precision=gpu.MIXED, members=[(1.0, g) for g in g2s]) as batch:
        batch.set_laplacian_form(True)
        views = [batch.member(m) for m in range(len(g2s))]
        for v, (ic, _) in zip(views, single):
            v.set_state(*ic)
            v.calc_metric()
        batch.symp4(dt, 3)
        got = [v.get_state() for v in views]
        for v in views:
            v.close()
    for (ic, ref), g in zip(single, got):
        for a, b in zip(g[:2], ref[:2]):
            assert np.array_equal(np.asarray(a), np.asarray(b))
        assert g[2] == ref[2] and g[3] == ref[3]


def test_separable_rejects_unsupported(gpu):
    with gpu.Lattice(128, nfld=2, precision=gpu.F64) as lat:
        with pytest.raises(gpu.LatticeError):
            lat.set_laplacian_form(True)
    with gpu.Lattice(64, nfld=2, precision=gpu.MIXED) as lat:
        with pytest.raises(gpu.LatticeError):
            lat.set_laplacian_form(True)
        lat.set_laplacian_form(False)
