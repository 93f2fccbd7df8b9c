"""The allocation-free entry point lat_spectrum_discrete_k_noalloc, through the
C ABI, against lat_spectrum_discrete_k on the same transforms: two calls in
a row on one GPU (the second with a smaller lattice, so the static scratch
holds the first call's larger contents), both kinds, bitwise equal (same
kernels, grid and order of additions).  Device buffers come from torch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc_vs_ctx(gpu, torch, n, seed):
    rng = np.random.default_rng(seed)
    Fk = np.fft.rfftn(rng.standard_normal((n, n, n)))
    ns = int(np.sqrt(3.0) * (n // 2) + 2)
    fk_dev = torch.from_numpy(np.ascontiguousarray(Fk).view(np.float64).copy()).to("cuda")
    s_dev = torch.full((ns,), np.nan, dtype=torch.float64, device="cuda")
    out = {}
    with gpu.Lattice(n, nfld=1, length=10.0, model=gpu.MODEL_PHI4, model_params=(1.0,)) as lat:
        for kind in (0, 1):
            ref = lat.spectrum_discrete_k(Fk, kind)
            torch.cuda.synchronize()
            lat.spectrum_discrete_k_noalloc(fk_dev.data_ptr(), s_dev.data_ptr(), kind)
            lat.synchronize()
            got = s_dev.cpu().numpy().copy()
            out[kind] = (got, ref)
    return out


def test_noalloc_spectrum_matches_allocating_entry(gpu):
    import torch

    print("case noalloc spectrum 64 then 32", flush=True)
    for n, seed in ((64, 1), (32, 2)):
        for kind, (got, ref) in _noalloc_vs_ctx(gpu, torch, n, seed).items():
            assert got.shape == ref.shape, (n, kind)
            assert np.array_equal(got, ref), (n, kind, np.abs(got - ref).max())


def test_noalloc_rejects_non_cubic(gpu):
    import torch

    with gpu.Lattice((16, 32, 16), nfld=1, length=10.0, model=gpu.MODEL_PHI4, model_params=(1.0,)) as lat:
        buf = torch.zeros(2 * 9 * 32 * 16, dtype=torch.float64, device="cuda")
        s = torch.zeros(32, dtype=torch.float64, device="cuda")
        with pytest.raises(gpu.LatticeError):
            lat.spectrum_discrete_k_noalloc(buf.data_ptr(), s.data_ptr(), 0)
