"""The learned corrector on MFMA: forward pass and gradients against a torch
(CPU, FP32) reference with the same BF16 rounding points, training on pairs
from the plain integrator, and the fallback to the plain integrator."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def bf16(x):
    """Round-to-nearest-even to BF16, returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def state(n, L, seed, amp=2e3):
    fld = np.empty((2, n, n, n))
    fldp = np.empty_like(fld)
    fld[0] = 1.0 + O.sample(n, L, -0.25, 1.0, amp, seed, 0)
    fldp[0] = -0.7 + O.sample(n, L, 0.25, 1.0, amp, seed, 1)
    fld[1] = O.sample(n, L, -0.25, 1.0, amp, seed, 2)
    fldp[1] = O.sample(n, L, 0.25, 1.0, amp, seed, 3)
    return fld, fldp


def random_layers(rng, scale=0.3):
    layers = []
    for cin, cout in [(8, 32), (32, 32), (32, 32), (32, 4)]:
        w = (rng.standard_normal((cout, 27, cin)) * scale / np.sqrt(27 * cin)).astype(np.float32)
        b = (rng.standard_normal(cout) * 0.1).astype(np.float32)
        layers.append((w, b))
    return layers


def torch_weight(w):
    """(cout, 27, cin) with tap = (dz+1)*9 + (dy+1)*3 + (dx+1) -> conv3d weight [cout][cin][dz][dy][dx]."""
    cout, _, cin = w.shape
    return torch.from_numpy(bf16(w).reshape(cout, 3, 3, 3, cin).transpose(0, 4, 1, 2, 3).copy())


def torch_forward(x0, layers, rounding=True):
    """x0: (nz, ny, nx, 8) BF16-valued; returns output (nz, ny, nx, 4) and the graph leaves."""
    F = torch.nn.functional
    h = torch.from_numpy(x0.transpose(3, 0, 1, 2)[None].copy())
    params = []
    for li, (w, b) in enumerate(layers):
        W = torch_weight(w).requires_grad_(True)
        B = torch.from_numpy(b.copy()).requires_grad_(True)
        params.append((W, B))
        z = F.conv3d(F.pad(h, (1, 1, 1, 1, 1, 1), mode="circular"), W, B)
        if li < 3:
            if rounding:
                z = z + (torch.from_numpy(bf16(z.detach().numpy())) - z).detach()
            a = z * torch.sigmoid(z)
            if rounding:
                a = a + (torch.from_numpy(bf16(a.detach().numpy())) - a).detach()
            h = a
        else:
            h = z
    return h[0].permute(1, 2, 3, 0), params


def packed_input(fld, fldp, g2):
    """Fluctuations about the per-channel spatial mean over the per-channel
    spatial standard deviation, then the uniform channels <phi>, <pi_phi>,
    log10(g^2/m^2) - 5 and log10(s_chi) / 10, in BF16."""
    v = np.stack([fld[0], fld[1], fldp[0], fldp[1]], axis=-1)
    m, sd = v.mean(axis=(0, 1, 2)), v.std(axis=(0, 1, 2))
    x = np.empty(v.shape[:3] + (8,), np.float32)
    x[..., :4] = (v - m) / sd
    x[..., 4], x[..., 5] = m[0], m[2]
    x[..., 6], x[..., 7] = np.log10(g2) - 5.0, np.log10(sd[1]) / 10.0
    return bf16(x)


def test_corrector_forward_matches_torch(gpu):
    print("case corrector forward", flush=True)
    n, L = 16, 10.0
    rng = np.random.default_rng(1)
    fld, fldp = state(n, L, 4)
    layers = random_layers(rng)
    with gpu.Lattice(n, nfld=2, length=L, model_params=(1.0, 1e4)) as lat, gpu.Corrector(3) as corr:
        lat.set_state(fld, fldp, 1.0, 0.0)
        for li, (w, b) in enumerate(layers):
            corr.set_layer(li, w, b)
        corr.set_norm(np.ones(4))
        out = corr.infer(lat)
        w0, b0 = corr.get_layer(0)
    assert np.array_equal(w0, layers[0][0]) and np.array_equal(b0, layers[0][1])
    ref, _ = torch_forward(packed_input(fld, fldp, 1e4), layers)
    ref = ref.detach().numpy()
    err = np.abs(out - ref)
    # same BF16 rounding points; only the FP32 summation order differs, which
    # can move a value across a BF16 rounding boundary now and then
    assert np.median(err) < 1e-3 * np.abs(ref).max()
    assert np.linalg.norm(out - ref) < 2e-2 * np.linalg.norm(ref)
    # a wrong MFMA operand layout or tap order gives O(1) errors
    assert np.corrcoef(out.ravel(), ref.ravel())[0, 1] > 0.9999


def test_corrector_gradients_match_torch(gpu):
    print("case corrector gradients", flush=True)
    n, L = 16, 10.0
    rng = np.random.default_rng(2)
    fld, fldp = state(n, L, 6)
    layers = random_layers(rng, scale=0.5)
    target = (rng.standard_normal((n, n, n, 4)) * 0.5).astype(np.float32)
    with gpu.Lattice(n, nfld=2, length=L, model_params=(1.0, 1e4)) as lat, gpu.Corrector(3) as corr:
        lat.set_state(fld, fldp, 1.0, 0.0)
        for li, (w, b) in enumerate(layers):
            corr.set_layer(li, w, b)
        corr.set_norm(np.ones(4))
        loss, grads = corr.gradients(lat, target)
    out, params = torch_forward(packed_input(fld, fldp, 1e4), layers)
    tl = ((out - torch.from_numpy(target)) ** 2).mean()
    tl.backward()
    assert abs(loss - tl.item()) < 1e-2 * tl.item()
    for li, ((W, B), (gw, gb)) in enumerate(zip(params, grads)):
        cout, _, cin = gw.shape
        ref_w = W.grad.numpy().transpose(0, 2, 3, 4, 1).reshape(cout, 27, cin)
        ref_b = B.grad.numpy()
        # BF16 gradient tensors between layers: a few per cent in norm
        assert np.linalg.norm(gw - ref_w) < 5e-2 * np.linalg.norm(ref_w), li
        assert np.linalg.norm(gb - ref_b) < 5e-2 * np.linalg.norm(ref_b), li
        cos = np.dot(gw.ravel(), ref_w.ravel()) / (np.linalg.norm(gw) * np.linalg.norm(ref_w))
        assert cos > 0.998, li


def test_corrector_training_and_fallback(gpu):
    """Training on integrator pairs lowers the loss; a zero threshold makes
    every step fall back to the plain integrator, reproducing Yoshida-4."""
    print("case corrector training", flush=True)
    n, L, dt, cf = 32, 10.0, 2e-3, 4
    fld, fldp = state(n, L, 8)
    kw = dict(nfld=2, length=L, model_params=(1.0, 1e3))
    with gpu.Lattice(n, **kw) as lat, gpu.Corrector(5) as corr:
        lat.set_state(fld, fldp, 1.0, 0.0)
        lat.calc_metric()
        losses = corr.train(lat, dt, cf, npairs=40, lr=2e-3)
        assert np.all(np.isfinite(losses))
        assert losses[-5:].mean() < 0.95 * losses[:5].mean()
        ratio = corr.get_norm()
        assert np.all(ratio > 0) and np.all(ratio < 1)
        start = lat.get_state()
        st = corr.step(lat, dt, cf, nsteps=2, threshold=0.0)
        # each step: the rejected coarse attempt (1 force evaluation) + cf Yoshida-4 steps
        assert st["fallbacks"] == 2 and st["force_evals"] == 2 * (1 + 3 * cf)
        got = lat.get_state()
        st2 = corr.step(lat, dt, cf, nsteps=1, threshold=1.0)
        assert st2["fallbacks"] == 0 and st2["force_evals"] == 1
        acc = lat.get_state()
    with gpu.Lattice(n, **kw) as ref:
        ref.set_state(start[0], start[1], start[2], start[3])
        ref.symp4(dt, 2 * cf)
        want = ref.get_state()
    assert np.linalg.norm(got[0] - want[0]) < 1e-12 * np.linalg.norm(want[0])
    assert np.linalg.norm(got[1] - want[1]) < 1e-11 * np.linalg.norm(want[1])
    # an accepted step = one leapfrog step of cf*dt plus a correction with no
    # homogeneous part: scale factor identical, channel means unchanged
    with gpu.Lattice(n, **kw) as ref:
        ref.set_state(got[0], got[1], got[2], got[3])
        ref.symp(2, cf * dt, 1)
        leap = ref.get_state()
    assert acc[2] == leap[2] and acc[3] == leap[3]
    for k in (0, 1):
        for i in (0, 1):
            d = acc[k][i] - leap[k][i]
            assert abs(d.mean()) < 1e-12 * (abs(leap[k][i].mean()) + leap[k][i].std())
            assert np.linalg.norm(d) > 0
