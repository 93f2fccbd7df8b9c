"""Device initial-condition sampler (sample(), evolve_spectral.f90:421-506 with a
Philox stream) against the oracle's restatement on the same seed/stream."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [16, 32, 48])
@pytest.mark.parametrize("gamma", [-0.25, 0.25])
def test_sample_matches_oracle(gpu, n, gamma):
    print("case sample", n, gamma, flush=True)
    L, m2eff, mpl, seed = 10.0, 2.5, 2e5, 1234
    with gpu.Lattice(n, nfld=2, length=L) as lat:
        lat.sample_field(1, 1 if gamma > 0 else 0, 0.0, gamma, m2eff, mpl, seed, 7)
        lat.sample_field(0, 0, 0.5, gamma, m2eff, mpl, seed, 7)
        f, p, _, _ = lat.get_state()
    got = (p if gamma > 0 else f)[1]
    ref = O.sample(n, L, gamma, m2eff, mpl, seed, 7)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    # the homogeneous value is added on top of the same sample
    assert np.abs(f[0] - (0.5 + ref)).max() <= 2e-16

