"""CPU pins of the warm-start network restatement (oracle/warmstart_net.py)."""
import numpy as np
import pytest
import scipy.signal
import torch

import oracle_import  # noqa: F401  (puts oracle/ on sys.path)
import warmstart_net as NR


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 7, 5))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    out = NR.conv3x3(x, w, b)
    for n in range(2):
        for o in range(4):
            ref = b[o] + sum(scipy.signal.correlate(x[n, i], w[o, i], mode="same") for i in range(3))
            assert np.allclose(out[n, o], ref, rtol=1e-13, atol=1e-13)


def test_pool_and_up_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 8, 6))
    y = rng.standard_normal((2, 3, 4, 3))
    assert np.isclose((NR.pool(x) * y).sum(), (x * NR.pool_bwd(y)).sum(), rtol=1e-13)
    assert np.isclose((NR.up(y) * x).sum(), (y * NR.up_bwd(x)).sum(), rtol=1e-13)
    assert np.allclose(NR.pool(NR.up(y)), y)


def test_conv_backward_is_the_adjoint():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 3, 6, 5))
    w = rng.standard_normal((4, 3, 3, 3))
    d = rng.standard_normal((2, 4, 6, 5))
    dx, dw, db = NR.conv3x3_bwd(x, w, d)
    zero = np.zeros(4)
    # <conv(x), d> is linear in x and in w
    assert np.isclose((NR.conv3x3(x, w, zero) * d).sum(), (dx * x).sum(), rtol=1e-12)
    assert np.isclose((NR.conv3x3(x, w, zero) * d).sum(), (dw * w).sum(), rtol=1e-12)
    assert np.allclose(db, d.sum(axis=(0, 2, 3)))


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    params = NR.init_params(seed=4)
    B, H, W = 2, 8, 8
    x = rng.standard_normal((B, 8, H, W))
    t = rng.standard_normal((B, 3, H, W))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    loss, grads, _ = NR.loss_and_grads(params, x, t)
    h = 1e-6
    for layer in range(8):
        for which in (0, 1):
            arr = params[layer][which]
            g = grads[layer][which]
            for _ in range(3):
                idx = tuple(rng.integers(0, s) for s in arr.shape)
                old = arr[idx]
                arr[idx] = old + h
                lp = NR.loss_and_grads(params, x, t)[0]
                arr[idx] = old - h
                lm = NR.loss_and_grads(params, x, t)[0]
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (layer, which, idx, fd, g[idx])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    a = np.concatenate([(rng.standard_normal(1000) * 10.0 ** rng.integers(-6, 6, 1000)).astype(np.float32),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0, 0.0])])
    ref = torch.from_numpy(a).to(torch.bfloat16).float().numpy()
    assert np.array_equal(NR.bf16_round(a), ref)


def test_adam_matches_torch():
    rng = np.random.default_rng(6)
    w0 = rng.standard_normal(10)
    p = torch.tensor(w0.copy(), requires_grad=True)
    opt = torch.optim.Adam([p], lr=1e-2)
    w, mom, vel = w0.copy(), np.zeros(10), np.zeros(10)
    for t in range(1, 6):
        g = rng.standard_normal(10)
        p.grad = torch.tensor(g.copy())
        opt.step()
        w, mom, vel = NR.adam_step(w, g, mom, vel, t, 1e-2)
    assert np.allclose(w, p.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_baseline_losses_closed_forms():
    """The untrained predictors' losses (sum |m - target|^2 / cells) on constructed data: targets all
    +z give 0 for the uniform +z and the best-uniform predictors; targets +z on half of each film and
    -z on the other half give 2 for both (the mean direction is 0, and every unit vector u then has
    loss 2 - 2 u . <target> = 2); m0 = target gives 0 for the m0 predictor."""
    import magnumaf.warmstart as WS
    n, s = 3, 8
    y = np.zeros((n, 3, s, s), np.float32)
    y[:, 2] = 1.0
    x = np.zeros((n, 8, s, s), np.float32)
    x[:, :3] = y
    b = WS.baseline_losses({"x": x, "y": y}, np.arange(n))
    assert b == {"m0": 0.0, "uniform_z": 0.0, "uniform_best": 0.0}
    y2 = y.copy()
    y2[:, 2, :, : s // 2] = -1.0
    x2 = x.copy()
    x2[:, 2] = 0.0
    x2[:, 0] = 1.0  # m0 along +x: |x - (+-z)|^2 = 2 everywhere
    b2 = WS.baseline_losses({"x": x2, "y": y2}, np.arange(n))
    assert abs(b2["m0"] - 2.0) < 1e-12 and abs(b2["uniform_z"] - 2.0) < 1e-12
    assert abs(b2["uniform_best"] - 2.0) < 1e-6


def test_bf16_bits_match_torch():
    """The BF16 weight file's rounding (round to nearest, ties to even) equals torch's float32 ->
    bfloat16 conversion, including ties and values near the float32 range's ends."""
    import magnumaf.warmstart as WS
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * np.float32(7.0),
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -0.0, 3.0e38, 1e-38], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(WS.bf16_bits(x), ref)
