"""CPU restatement (numpy, FP64 or FP32) of the learned warm-start network.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU leg use this file.
The network has no counterpart in the reference (magnum.af has no learned component); it is the
checker for the HIP forward/backward kernels of magnum.af_amd/csrc/net.hip, which implement the
same operations in the same order.

Network (a 4-level fully-convolutional encoder-decoder):
    x [B, 8, H, W]: m0 (3 planes) + 5 broadcast parameter planes, H and W multiples of 8
    e1 = relu(conv(x,                  W1 [32, 8]))        H
    e2 = relu(conv(pool(e1),           W2 [64, 32]))       H/2
    e3 = relu(conv(pool(e2),           W3 [128, 64]))      H/4
    e4 = relu(conv(pool(e3),           W4 [128, 128]))     H/8
    d3 = relu(conv(cat(up(e4), e3),    W5 [64, 256]))      H/4
    d2 = relu(conv(cat(up(d3), e2),    W6 [32, 128]))      H/2
    d1 = relu(conv(cat(up(d2), e1),    W7 [32, 64]))       H
    y  =      conv(d1,                 W8 [3, 32])         H
    m  = y / sqrt(|y|^2 + eps)          per cell
    loss = sum |m - target|^2 / (B H W)
conv: 3x3, stride 1, zero padding 1, plus a per-channel bias; pool: 2x2 average; up: nearest x2;
cat: channel concatenation (first operand first).

Parity pins (tests/test_warmstart_cpu.py): conv against scipy.signal.correlate (a textbook
routine), and every gradient against central finite differences of this file's own loss on tiny
shapes, so a dropped term, a wrong flip of the 3x3 kernel or a transposed operand fails.
"""
import numpy as np

# (in_channels, out_channels) of the eight convolutions, in order
LAYERS = [(8, 32), (32, 64), (64, 128), (128, 128), (256, 64), (128, 32), (64, 32), (32, 3)]
EPS = 1e-12


def conv3x3(x, w, b):
    """x [B, Ci, H, W], w [Co, Ci, 3, 3], b [Co] -> [B, Co, H, W] (cross-correlation, zero pad 1)."""
    B, Ci, H, W = x.shape
    Co = w.shape[0]
    xp = np.zeros((B, Ci, H + 2, W + 2), dtype=x.dtype)
    xp[:, :, 1:-1, 1:-1] = x
    out = np.zeros((B, Co, H, W), dtype=x.dtype)
    for ky in range(3):
        for kx in range(3):
            patch = xp[:, :, ky:ky + H, kx:kx + W]  # [B, Ci, H, W]
            out += np.einsum("bihw,oi->bohw", patch, w[:, :, ky, kx], optimize=True)
    return out + b[None, :, None, None]


def conv3x3_bwd(x, w, dout):
    """Gradients of conv3x3 with respect to x, w and b."""
    B, Ci, H, W = x.shape
    xp = np.zeros((B, Ci, H + 2, W + 2), dtype=x.dtype)
    xp[:, :, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for ky in range(3):
        for kx in range(3):
            patch = xp[:, :, ky:ky + H, kx:kx + W]
            dw[:, :, ky, kx] = np.einsum("bohw,bihw->oi", dout, patch, optimize=True)
            dxp[:, :, ky:ky + H, kx:kx + W] += np.einsum("bohw,oi->bihw", dout, w[:, :, ky, kx], optimize=True)
    db = dout.sum(axis=(0, 2, 3))
    return dxp[:, :, 1:-1, 1:-1], dw, db


def pool(x):
    B, C, H, W = x.shape
    return x.reshape(B, C, H // 2, 2, W // 2, 2).mean(axis=(3, 5))


def pool_bwd(d):
    return np.repeat(np.repeat(d, 2, axis=2), 2, axis=3) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)


def up_bwd(d):
    B, C, H, W = d.shape
    return d.reshape(B, C, H // 2, 2, W // 2, 2).sum(axis=(3, 5))


def init_params(seed=0, dtype=np.float64):
    """He-normal weights, zero biases; the same generator the library's mafx_net_create uses is not
    required: tests hand these arrays to the library."""
    rng = np.random.default_rng(seed)
    ps = []
    for ci, co in LAYERS:
        w = rng.standard_normal((co, ci, 3, 3)) * np.sqrt(2.0 / (9 * ci))
        ps.append((w.astype(dtype), np.zeros(co, dtype=dtype)))
    return ps


def forward(params, x):
    """Returns (m, cache)."""
    (w1, b1), (w2, b2), (w3, b3), (w4, b4), (w5, b5), (w6, b6), (w7, b7), (w8, b8) = params
    c = {"x": x}
    c["e1"] = np.maximum(conv3x3(x, w1, b1), 0)
    c["p1"] = pool(c["e1"])
    c["e2"] = np.maximum(conv3x3(c["p1"], w2, b2), 0)
    c["p2"] = pool(c["e2"])
    c["e3"] = np.maximum(conv3x3(c["p2"], w3, b3), 0)
    c["p3"] = pool(c["e3"])
    c["e4"] = np.maximum(conv3x3(c["p3"], w4, b4), 0)
    c["c3"] = np.concatenate([up(c["e4"]), c["e3"]], axis=1)
    c["d3"] = np.maximum(conv3x3(c["c3"], w5, b5), 0)
    c["c2"] = np.concatenate([up(c["d3"]), c["e2"]], axis=1)
    c["d2"] = np.maximum(conv3x3(c["c2"], w6, b6), 0)
    c["c1"] = np.concatenate([up(c["d2"]), c["e1"]], axis=1)
    c["d1"] = np.maximum(conv3x3(c["c1"], w7, b7), 0)
    c["y"] = conv3x3(c["d1"], w8, b8)
    c["n"] = np.sqrt((c["y"] ** 2).sum(axis=1, keepdims=True) + EPS)
    m = c["y"] / c["n"]
    return m, c


def loss_and_grads(params, x, target):
    """loss = sum |m - target|^2 / (B H W) and its gradients [(dW, db)] in layer order."""
    m, c = forward(params, x)
    B, _, H, W = x.shape
    scale = 1.0 / (B * H * W)
    loss = ((m - target) ** 2).sum() * scale
    dm = 2.0 * (m - target) * scale
    # m = y / n: dy = (dm - m (m . dm)) / n
    dy = (dm - m * (m * dm).sum(axis=1, keepdims=True)) / c["n"]
    ws = [p[0] for p in params]
    grads = [None] * 8
    dd1, grads[7] = _conv_bwd(c["d1"], ws[7], dy)
    dd1 = dd1 * (c["d1"] > 0)
    dc1, grads[6] = _conv_bwd(c["c1"], ws[6], dd1)
    n_up = c["d2"].shape[1]
    dd2 = up_bwd(dc1[:, :n_up]) * (c["d2"] > 0)
    de1_skip = dc1[:, n_up:]
    dc2, grads[5] = _conv_bwd(c["c2"], ws[5], dd2)
    n_up = c["d3"].shape[1]
    dd3 = up_bwd(dc2[:, :n_up]) * (c["d3"] > 0)
    de2_skip = dc2[:, n_up:]
    dc3, grads[4] = _conv_bwd(c["c3"], ws[4], dd3)
    n_up = c["e4"].shape[1]
    de4 = up_bwd(dc3[:, :n_up]) * (c["e4"] > 0)
    de3_skip = dc3[:, n_up:]
    dp3, grads[3] = _conv_bwd(c["p3"], ws[3], de4)
    de3 = (pool_bwd(dp3) + de3_skip) * (c["e3"] > 0)
    dp2, grads[2] = _conv_bwd(c["p2"], ws[2], de3)
    de2 = (pool_bwd(dp2) + de2_skip) * (c["e2"] > 0)
    dp1, grads[1] = _conv_bwd(c["p1"], ws[1], de2)
    de1 = (pool_bwd(dp1) + de1_skip) * (c["e1"] > 0)
    _, grads[0] = _conv_bwd(c["x"], ws[0], de1)
    return loss, grads, m


def _conv_bwd(x, w, d):
    dx, dw, db = conv3x3_bwd(x, w, d)
    return dx, (dw, db)


def bf16_round(a):
    """Round-to-nearest-even to bfloat16, returned as float32 (the weights the kernels use)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def adam_step(w, g, mom, vel, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma & Ba 2015) with bias correction; returns the updated (w, mom, vel)."""
    mom = b1 * mom + (1 - b1) * g
    vel = b2 * vel + (1 - b2) * g * g
    mh = mom / (1 - b1 ** t)
    vh = vel / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), mom, vel
