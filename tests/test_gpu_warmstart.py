"""GPU parity of the warm-start network kernels (csrc/net.hip) against the numpy restatement
oracle/warmstart_net.py, evaluated at the BF16-rounded weights the kernels use.  Tolerances: FP32
accumulation of up to 9 * 256 products per output."""
import numpy as np
import pytest

import oracle_import  # noqa: F401
import warmstart_net as NR

pytestmark = pytest.mark.gpu


def _setup(B, H, W, seed):
    import torch
    import magnumaf.warmstart as WS
    rng = np.random.default_rng(seed)
    params = NR.init_params(seed=seed)
    for i, (w, b) in enumerate(params):  # non-zero biases exercise the bias path
        params[i] = (w, rng.standard_normal(b.shape) * 0.1)
    flat = np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in params]).astype(np.float32)
    net = WS.WarmStartNet()
    net.set_params(flat)
    # what the kernels compute with: BF16 weights, FP32 biases
    used = [(NR.bf16_round(w.astype(np.float32)).astype(np.float64), b.astype(np.float32).astype(np.float64))
            for w, b in params]
    x = rng.standard_normal((B, 8, H, W)).astype(np.float32)
    t = rng.standard_normal((B, 3, H, W))
    t = (t / np.linalg.norm(t, axis=1, keepdims=True)).astype(np.float32)
    return net, used, x, t, torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()


@pytest.mark.parametrize("shape", [(2, 16, 16), (1, 24, 40), (3, 8, 8)])
def test_forward_matches_oracle(shape):
    print(f"\n[net forward] B,H,W={shape}", flush=True)
    net, used, x, t, xd, td = _setup(*shape, seed=1)
    m = net.forward(xd).cpu().numpy()
    ref, _ = NR.forward(used, x.astype(np.float64))
    assert np.max(np.abs(m - ref)) <= 2e-4, np.max(np.abs(m - ref))


@pytest.mark.parametrize("shape", [(2, 16, 16), (1, 24, 40)])
def test_gradients_match_oracle(shape):
    print(f"\n[net gradients] B,H,W={shape}", flush=True)
    net, used, x, t, xd, td = _setup(*shape, seed=2)
    loss, g = net.gradients(xd, td)
    ref_loss, ref_grads, _ = NR.loss_and_grads(used, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - ref_loss) <= 1e-5 * abs(ref_loss)
    off = 0
    for layer, (dw, db) in enumerate(ref_grads):
        for ref in (dw, db):
            got = g[off:off + ref.size].reshape(ref.shape)
            off += ref.size
            err = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
            assert err <= 1e-3, (layer, ref.shape, err)
    assert off == g.size


def test_training_reduces_loss():
    print("\n[net training] B=4 16x16, 40 Adam steps", flush=True)
    net, used, x, t, xd, td = _setup(4, 16, 16, seed=3)
    first = net.train_step(xd, td, 2e-3)
    for _ in range(39):
        last = net.train_step(xd, td, 2e-3)
    assert last < 0.7 * first, (first, last)


def test_warm_start_entry_gives_unit_vectors():
    import magnumaf as M
    import magnumaf.warmstart as WS
    print("\n[warm start] 16x16 film", flush=True)
    net = WS.WarmStartNet()
    rng = np.random.default_rng(4)
    mesh = M.Mesh(16, 16, 1, 3e-9, 3e-9, 3e-9)
    state = M.State(mesh, 8e5, WS.random_m(rng, 16, 16))
    net.warm_start(state, WS.random_params(rng))
    m = state.m
    assert np.all(np.isfinite(m))
    assert np.allclose(np.linalg.norm(m, axis=-1), 1.0, atol=1e-12)


WEIGHTS = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "warmstart_weights_bf16.npy")


def test_trained_warm_start_gain_at_equal_energy():
    """The committed network (tests/golden/warmstart_weights_bf16.npy, from `bench.py --config C4
    --c4-weights ...`: trained on 4096 relaxed 64x64 films, the best held-out epoch, stored as BF16)
    against plain LBFGS on 64x64 films from seeds the training never saw, both to
    max|m x H| / max|H| <= 1e-6.  Equal final energies decide something only where the film's
    minimum is defined by its start: plain LBFGS from m0 and from m0 + 1e-6 noise end within 1e-9 of
    each other.  Where that holds and the warm start reaches the same energy (within 1e-9), it must
    take at most half the iterations of plain LBFGS (median), on at least three films."""
    import magnumaf as M
    import magnumaf.warmstart as WS
    net = WS.WarmStartNet.load(WEIGHTS)
    mesh = M.Mesh(64, 64, 1, WS.CELL, WS.CELL, WS.CELL)
    recs = []
    for i in range(10):
        rng = np.random.default_rng([4242, i])
        p = WS.random_params(rng)
        m0 = WS.random_m(rng, 64, 64)
        mp = m0 + 1e-6 * np.random.default_rng([4242, i, 1]).standard_normal(m0.shape)
        mp /= np.linalg.norm(mp, axis=-1, keepdims=True)
        _, sp, _ = WS.relax(mesh, m0, p, 1e-6, 20000)
        _, sq, _ = WS.relax(mesh, mp, p, 1e-6, 20000)
        _, sw, _ = WS.relax(mesh, m0, p, 1e-6, 20000, net=net)
        rel = lambda a, b: abs(a - b) / abs(b)
        r = dict(film=i, it_plain=sp["iterations"], it_warm=sw["iterations"], defined=rel(sq["E"], sp["E"]) <= 1e-9,
                 equal=rel(sw["E"], sp["E"]) <= 1e-9, dE_warm=rel(sw["E"], sp["E"]), warm_lower=sw["E"] < sp["E"],
                 stops=(sp["stop_reason"], sw["stop_reason"]))
        recs.append(r)
        print(f"\n[warm start 64x64] {r}", flush=True)
    both = [r for r in recs if r["defined"] and r["equal"]]
    ratio = [r["it_plain"] / r["it_warm"] for r in both]
    print(f"[warm start 64x64] {len(both)} films with a defined minimum reached by both; iteration ratios "
          f"{[round(x, 2) for x in ratio]}", flush=True)
    assert all(r["stops"] == (3, 3) for r in recs)
    assert len(both) >= 3
    assert np.median(ratio) >= 2.0
