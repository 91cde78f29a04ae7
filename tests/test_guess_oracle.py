"""Pins of the learned guess's CPU oracle (oracle/guess_oracle.py), the checker of the
library's training kernels: the analytic gradient against central finite
differences of the loss, and against torch autograd of an independent
implementation of the same network (oracle/guess_torch.py, build_model); the
BF16-storage forward against the torch restatement of the library's inference;
Adam against torch.optim.Adam. CPU only.
"""
import math

import numpy as np
import pytest

import guess_oracle as G


def smooth_pair(n, B, seed):
    """B pairs of smooth periodic FP64 pressures (p_prev, p): a few random Fourier
    modes, p a perturbation of p_prev."""
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    th = 2 * np.pi / n

    def field():
        f = np.zeros((n, n))
        for _ in range(6):
            kx, ky = rng.integers(-3, 4, 2)
            f += rng.normal() * np.cos(th * (kx * x + ky * y) + rng.uniform(0, 2 * np.pi))
        return f

    pp = np.stack([field() for _ in range(B)])
    p = pp + 0.05 * np.stack([field() for _ in range(B)])
    return p, pp


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_layout_matches_the_blob_format():
    from fluidsim.guess_net import LAYERS, load_blob
    lay, total = G.layout()
    assert [tuple(c) for c in G.LAYERS[:14]] == [tuple(c) for c in LAYERS]
    theta = G.init_params(3)
    layers = load_blob(G.to_blob(theta))
    for (w, b), (wo, bo, ci, co, k) in zip(layers, lay):
        # torch layout [cout][cin][k][k] back to the blob's [cout][k*k][cin]
        np.testing.assert_array_equal(w.transpose(0, 2, 3, 1).reshape(-1), theta[wo:bo].astype(np.float32))
        np.testing.assert_array_equal(b, theta[bo:bo + co].astype(np.float32))
    assert total == sum(co * k * k * ci + co for (ci, co), k in zip(G.LAYERS, G.KSIZE))


def test_gradient_matches_central_finite_differences():
    # brute force on a tiny grid: every layer's weights and bias, a few entries each
    n, B, h = 8, 2, 2 * math.pi / 8
    p, pp = smooth_pair(n, B, 0)
    x, t, _ = G.make_inputs(p, pp, h)
    theta = G.init_params(7)
    loss, grad = G.loss_and_gradient(theta, x, t)
    assert abs(loss - G.loss_only(theta, x, t)) <= 1e-14 * abs(loss)
    rng = np.random.default_rng(1)
    lay, _ = G.layout()
    picks = []
    for w, b, ci, co, k in lay:
        picks += list(rng.integers(w, b, 3)) + [b + int(rng.integers(0, co))]
    eps = 1e-6
    num, ana = [], []
    for i in picks:
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += eps
        tm[i] -= eps
        num.append((G.loss_only(tp, x, t) - G.loss_only(tm, x, t)) / (2 * eps))
        ana.append(grad[i])
    num, ana = np.array(num), np.array(ana)
    assert np.max(np.abs(ana)) > 0
    # each entry to the finite difference's own truncation/rounding error
    np.testing.assert_allclose(ana, num, rtol=1e-5, atol=1e-9 * np.max(np.abs(ana)))


def test_gradient_matches_torch_autograd():
    # an independent implementation (torch conv2d / avg_pool2d / interpolate, NCHW)
    torch = pytest.importorskip("torch")
    from guess_torch import model_from_blob
    n, B, h = 16, 2, 2 * math.pi / 16
    p, pp = smooth_pair(n, B, 2)
    x, t, _ = G.make_inputs(p, pp, h)
    theta = G.init_params(11)
    # the blob is float32: use the float32-exact parameters in both
    theta = theta.astype(np.float32).astype(np.float64)
    loss, grad = G.loss_and_gradient(theta, x, t)
    model = model_from_blob(G.to_blob(theta), torch.float64)
    xt = torch.from_numpy(x.transpose(0, 3, 1, 2).copy())
    tt = torch.from_numpy(t[:, None].copy())
    lt = torch.mean((model(xt) - tt) ** 2)
    lt.backward()
    assert abs(float(lt) - loss) <= 1e-12 * loss
    mods = list(model.convs) + [model.out]
    lay, _ = G.layout()
    for l, (m, (w, b, ci, co, k)) in enumerate(zip(mods, lay)):
        gw = m.weight.grad.numpy().transpose(0, 2, 3, 1).reshape(-1)
        gb = m.bias.grad.numpy()
        assert rel(grad[w:b], gw) < 1e-10, l
        assert rel(grad[b:b + co], gb) < 1e-10, l


def test_bf16_forward_matches_the_inference_restatement():
    # the storage-point emulation against guess_torch.reference_guess (torch,
    # float64, the inference's BF16 roundings)
    pytest.importorskip("torch")
    from guess_torch import reference_guess
    n, h = 16, 2 * math.pi / 16
    p, pp = smooth_pair(n, 1, 4)
    x, t, sc = G.make_inputs(p, pp, h)
    theta = G.init_params(5)
    _, _, o = G.forward(theta, x, bf16_storage=True)
    r = (G.laplacian(p, h) - G.laplacian(pp, h))[0]
    ref = reference_guess(G.to_blob(theta), r.astype(np.float32).reshape(-1), pp[0].reshape(-1), n, h)
    o_ref = (ref - pp[0].reshape(-1)) / (sc[0, 0] * h * h)
    assert rel(o[0].reshape(-1), o_ref) < 1e-6


def test_bf16_gradient_tracks_the_exact_gradient():
    # the rounding points move the gradient by BF16-sized amounts, no more: a dropped
    # term or a misplaced mask would show as an O(1) difference
    n, B, h = 16, 2, 2 * math.pi / 16
    p, pp = smooth_pair(n, B, 6)
    x, t, _ = G.make_inputs(p, pp, h)
    theta = G.init_params(9)
    _, g64 = G.loss_and_gradient(theta, x, t)
    _, g16 = G.loss_and_gradient(theta, x, t, bf16_storage=True)
    lay, _ = G.layout()
    for l, (w, b, ci, co, k) in enumerate(lay):
        assert rel(g16[w:b + co], g64[w:b + co]) < 0.05, l


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    th = rng.normal(size=50).astype(np.float32)
    w = torch.nn.Parameter(torch.from_numpy(th.copy()))
    opt = torch.optim.Adam([w], lr=1e-3)
    ours, m1, m2 = th.astype(np.float64), np.zeros(50), np.zeros(50)
    for step in range(1, 6):
        g = rng.normal(size=50).astype(np.float32)
        w.grad = torch.from_numpy(g.copy())
        opt.step()
        ours, m1, m2 = G.adam(ours, g, m1, m2, step, 1e-3)
    np.testing.assert_allclose(ours, w.detach().numpy(), rtol=0, atol=2e-7)
