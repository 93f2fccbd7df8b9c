"""The field force pipeline in the blocked k-space row order (the rows of 64
consecutive kz planes of one ky adjacent; cubic 256 and 1024) against the
same pipeline in FFTW order (lat_set_kspace_order(ctx, 0)) from identical
initial conditions.  Every column transform does the same arithmetic on the
same values in both orders; what differs is which sites each workgroup of
the x C2R + kick sweep reduces, so the reduction partials (KE, GE, PE) and
through them the FRW background and the kick coefficients agree to FP64
rounding only.  Bars: FP64 form 1e-12 relative on fields, momenta, a, p_a,
energies and spectra; mixed form the same on a, p_a and the energies, and
the FP32 rounding of the fluctuations (1e-6) on fields, momenta and
spectra.  The pipeline itself is held to the oracle at 256^3 in
test_gpu_long_parity."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PHI0 = 1.0093430384226378929425913902459
DPHI0 = -0.7137133070120812430962278466136


def _run(la, n, prec, blocked, steps):
    with la.Lattice(n, nfld=2, length=10.0, precision=prec, model_params=(1.0, 2.5e5)) as lat:
        lat.set_kspace_order(blocked)
        m2 = [1.0, 2.5e5 * PHI0 ** 2]
        lat.sample_field(0, 0, PHI0, -0.25, m2[0], 2e5, 20240601, 0)
        lat.sample_field(0, 1, DPHI0, 0.25, m2[0], 2e5, 20240601, 1)
        lat.sample_field(1, 0, 0.0, -0.25, m2[1], 2e5, 20240601, 2)
        lat.sample_field(1, 1, 0.0, 0.25, m2[1], 2e5, 20240601, 3)
        lat.calc_metric()
        lat.symp4((10.0 / n) / 40.0, steps)
        e = lat.energies()
        spec = lat.output_spectra(1)
        f, p, a, pa = lat.get_state()
        # a strided subsample of the fluctuations about each field's mean
        # (the 1024^3 arrays are 17 GB each)
        sub = (slice(None), slice(None, None, 7), slice(None, None, 5), slice(None, None, 3))
        fm, pm = f.mean(axis=(1, 2, 3)), p.mean(axis=(1, 2, 3))
        fs = f[sub] - fm[:, None, None, None]
        ps = p[sub] - pm[:, None, None, None]
        del f, p
    return fs, ps, a, pa, e, spec


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n,precision,steps", [(256, "f64", 3), (256, "mixed", 3), (1024, "mixed", 1)],
                         ids=["256-f64", "256-mixed", "1024-mixed"])
def test_blocked_order_matches_fftw_order(gpu, n, precision, steps):
    print("case kspace order", n, precision, steps, flush=True)
    prec = gpu.F64 if precision == "f64" else gpu.MIXED
    fs, ps, a, pa, e, spec = _run(gpu, n, prec, True, steps)
    fs0, ps0, a0, pa0, e0, spec0 = _run(gpu, n, prec, False, steps)
    df, dp = _rel(fs, fs0), _rel(ps, ps0)
    ds = max(float(np.max(np.abs(spec[q] - spec0[q])) / np.max(np.abs(spec0[q]))) for q in range(4))
    de = max(abs(e[k] - e0[k]) / abs(e0[k]) for k in ("rho", "ke", "pe", "ge"))
    print(n, precision, "fluctuations: fields %.2e momenta %.2e; a %.2e p_a %.2e spectra %.2e energies %.2e"
          % (df, dp, abs(a - a0) / abs(a0), abs(pa - pa0) / abs(pa0), ds, de), flush=True)
    assert abs(a - a0) <= 1e-12 * abs(a0) and abs(pa - pa0) <= 1e-12 * abs(pa0)
    assert de <= 1e-12
    bar = 1e-12 if precision == "f64" else 1e-6
    assert df <= bar and dp <= bar and ds <= bar
