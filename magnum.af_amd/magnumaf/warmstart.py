"""Learned warm start for the LBFGS minimizer (no counterpart in the reference).

A convolutional encoder-decoder (libmafx `mafx_net_*`, kernels in csrc/net.hip) maps an initial
film magnetisation m0 plus five material/field parameters to a predicted relaxed state; the plain
FP64 LBFGS of the library (`LBFGS_Minimizer`) then refines that prediction to the torque
tolerance.  Plain LBFGS from the raw initial state is both the fallback and the judge of the final
energy.

    net = WarmStartNet()                       # He-normal weights, or WarmStartNet.load(path)
    data = generate_dataset(4096, seed=1)      # FP64 LBFGS relaxations of 64x64x1 films
    train(net, data, epochs=20)                # forward/backward/Adam on the GPU
    evaluate(net, n=8, size=256)               # warm start + LBFGS vs plain LBFGS

Films: nx = ny = size cells of 3 nm cubes, one layer.  Terms (the dataset's choice, recorded with
every sample): demag, exchange A, interface DMI D (axis z), uniaxial anisotropy Ku (axis z) and
a perpendicular field Bz.  Parameter ranges: Ms in [5e5, 1.2e6] A/m, A in [1e-11, 2e-11] J/m,
D in [1e-3, 4e-3] J/m^2, Ku in [2e5, 9e5] J/m^3, Bz in [0, 0.5] T.
"""
import ctypes as C
import math
import sys
import threading
import time

import numpy as np

from . import (DemagField, DmiField, ExchangeField, ExternalField, LBFGS_Minimizer, Mesh, State,
               UniaxialAnisotropyField, _abi)
from ._abi import check, lib

MU0 = 4e-7 * math.pi
RANGES = (("Ms", 5e5, 1.2e6), ("A", 1e-11, 2e-11), ("D", 1e-3, 4e-3), ("Ku", 2e5, 9e5), ("Bz", 0.0, 0.5))
CELL = 3e-9


def bf16_bits(x):
    """float32 -> BF16 bit patterns (uint16), round to nearest, ties to even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the warm-start network needs a GPU")
    return torch


def planes(p):
    """The five input planes: each parameter mapped linearly from its range onto [-1, 1]."""
    return np.array([2.0 * (p[name] - lo) / (hi - lo) - 1.0 for name, lo, hi in RANGES], dtype=np.float32)


def film_terms(mesh, p):
    return [DemagField(mesh), ExchangeField(p["A"]), DmiField(p["D"], (0.0, 0.0, 1.0)),
            UniaxialAnisotropyField(p["Ku"], (0.0, 0.0, 1.0)),
            ExternalField(homogeneous=(0.0, 0.0, p["Bz"] / MU0))]


def random_params(rng):
    return {name: float(rng.uniform(lo, hi)) for name, lo, hi in RANGES}


def random_m(rng, nx, ny):
    m = rng.standard_normal((nx, ny, 1, 3))
    return m / np.linalg.norm(m, axis=-1, keepdims=True)


class WarmStartNet:
    """Owns one libmafx network (parameters, Adam state, activations) on torch's current stream."""

    def __init__(self):
        torch = _torch()
        torch.cuda.init()
        h = C.c_void_p()
        check(lib().mafx_net_create(C.c_void_p(torch.cuda.current_stream().cuda_stream), C.byref(h)))
        self.h = h
        n = C.c_int64()
        check(lib().mafx_net_param_count(self.h, C.byref(n)))
        self.n_params = n.value

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the process releases the device memory itself
        if getattr(self, "h", None) and not _finalizing():
            lib().mafx_net_destroy(self.h)
            self.h = None

    def get_params(self):
        out = np.empty(self.n_params, dtype=np.float32)
        check(lib().mafx_net_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        check(lib().mafx_net_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def save(self, path):
        np.save(path, self.get_params())

    def save_bf16(self, path):
        """The parameters as BF16 bit patterns (uint16, round to nearest even): half the file of
        save(), and the same network, since the kernels compute with the BF16 rounding of the
        weights.  (Biases are rounded too; the kernels keep them in FP32, so a network loaded from
        this file is the BF16-rounded one.)"""
        np.save(path, bf16_bits(self.get_params()))

    @classmethod
    def load(cls, path):
        """From save() (float32) or save_bf16() (uint16 BF16 bits)."""
        net = cls()
        p = np.load(path, allow_pickle=False)
        if p.dtype == np.uint16:
            p = (p.astype(np.uint32) << 16).view(np.float32)
        net.set_params(p)
        return net

    def forward(self, x):
        """x: torch float32 [B, 8, H, W] on the GPU -> unit vectors [B, 3, H, W]."""
        torch = _torch()
        x = x.contiguous()
        B, _, H, W = x.shape
        m = torch.empty((B, 3, H, W), dtype=torch.float32, device=x.device)
        check(lib().mafx_net_forward(self.h, C.c_void_p(x.data_ptr()), B, H, W, C.c_void_p(m.data_ptr())))
        return m

    def gradients(self, x, target):
        x, target = x.contiguous(), target.contiguous()
        B, _, H, W = x.shape
        g = np.empty(self.n_params, dtype=np.float32)
        loss = C.c_double()
        check(lib().mafx_net_gradients(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(target.data_ptr()), B, H, W,
                                       g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return loss.value, g

    def train_step(self, x, target, lr):
        B, _, H, W = x.shape
        loss = C.c_double()
        check(lib().mafx_net_train_step(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(target.data_ptr()), B, H, W,
                                        C.c_float(lr), C.byref(loss)))
        return loss.value

    def warm_start(self, state, p):
        """Overwrites state.m (a film, nz = 1) with the network's prediction from the current m."""
        torch = _torch()
        mesh = state.mesh
        out = torch.empty_like(state.m_dev)
        pl = planes(p)
        check(lib().mafx_net_warm_start(self.h, C.byref(mesh._c()), C.c_void_p(state.m_dev.data_ptr()),
                                        pl.ctypes.data_as(C.POINTER(C.c_float)), C.c_void_p(out.data_ptr())))
        state.m_dev.copy_(out)


def warm_start_layers(net, state, p):
    """The network's prediction, z-plane by z-plane, for a state of nz >= 1 planes: each plane is
    fed to the network as a film with the parameters p (the network has no input for couplings
    between planes or for in-plane fields), and the predicted unit vectors overwrite state.m."""
    torch = _torch()
    mesh = state.mesh
    nx, ny, nz = mesh.nx, mesh.ny, mesh.nz
    m = state.m_dev.view(3, nz, ny, nx)
    pl = torch.from_numpy(planes(p)).to(m.device)
    x = torch.empty((1, 8, ny, nx), dtype=torch.float32, device=m.device)
    for z in range(nz):
        x[0, :3] = m[:, z].float()
        x[0, 3:] = pl[:, None, None]
        m[:, z] = net.forward(x)[0].double()


def input_planes(m0, p):
    """x [8, ny, nx] float32 from m0 (nx, ny, 1, 3) and the parameter dict."""
    nx, ny = m0.shape[0], m0.shape[1]
    x = np.empty((8, ny, nx), dtype=np.float32)
    x[:3] = m0[:, :, 0, :].transpose(2, 1, 0)
    x[3:] = planes(p)[:, None, None]
    return x


def relax(mesh, m0, p, torque_tol, maxiter, net=None):
    """LBFGS (FP64, the library's) from m0, or from the network's prediction when net is given.
    Returns (state, stats, seconds); the seconds include the network inference."""
    torch = _torch()
    terms = film_terms(mesh, p)
    state = State(mesh, p["Ms"], m0)
    state._context(terms)  # tensor and plans outside the timed region
    mini = LBFGS_Minimizer(terms, tol=1e-20, maxiter=maxiter, torque_tol=torque_tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if net is not None:
        net.warm_start(state, p)
    E = mini.minimize(state)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return state, dict(mini.last_stats, E=E), dt


def generate_dataset(n, seed=1, size=64, torque_tol=1e-4, maxiter=20000, threads=8, log=None):
    """n relaxed samples: x [n, 8, size, size] and y [n, 3, size, size] (float32), with the
    parameters, energies and LBFGS statistics.  Samples are relaxed concurrently by `threads`
    host threads, each with its own stream and contexts."""
    torch = _torch()
    X = np.empty((n, 8, size, size), dtype=np.float32)
    Y = np.empty((n, 3, size, size), dtype=np.float32)
    meta = [None] * n
    mesh = Mesh(size, size, 1, CELL, CELL, CELL)
    errors = []

    def work(tid):
        try:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                for i in range(tid, n, threads):
                    rng = np.random.default_rng([seed, i])
                    p = random_params(rng)
                    m0 = random_m(rng, size, size)
                    state, st, dt = relax(mesh, m0, p, torque_tol, maxiter)
                    X[i] = input_planes(m0, p)
                    Y[i] = state.m[:, :, 0, :].transpose(2, 1, 0)
                    meta[i] = dict(p, E=st["E"], iterations=st["iterations"], stop=st["stop_reason"],
                                   torque=st["final_torque_rel"], seconds=dt)
                    if log and i % 256 == 0:
                        log(f"[dataset] sample {i}: {st['iterations']} iterations, {dt:.2f} s")
        except Exception as e:  # reported after the join
            errors.append(e)

    ts = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    if errors:
        raise errors[0]
    return {"x": X, "y": Y, "meta": meta, "seed": seed, "size": size, "torque_tol": torque_tol}


def split(n, holdout=0.1, seed=0):
    """(held-out, training) sample indices of train()'s split."""
    idx = np.random.default_rng(seed).permutation(n)
    nh = int(n * holdout)
    return idx[:nh], idx[nh:]


def baseline_losses(data, idx):
    """The network's loss, sum |m - target|^2 / cells, for three predictors that need no training,
    over the samples idx: m0 itself; the uniform state +z; and, per sample, the uniform state along
    that sample's own mean target direction (the best any uniform prediction can do)."""
    x, y = data["x"][idx].astype(np.float64), data["y"][idx].astype(np.float64)
    cells = y.shape[0] * y.shape[2] * y.shape[3]
    m0 = x[:, :3]
    zhat = np.zeros_like(y)
    zhat[:, 2] = 1.0
    mean = y.mean(axis=(2, 3), keepdims=True)
    nrm = np.linalg.norm(mean, axis=1, keepdims=True)
    # a zero mean leaves every unit vector equally good (loss 2); +z stands for them
    best = np.where(nrm > 0, mean / np.where(nrm > 0, nrm, 1.0), zhat[:, :, :1, :1])
    return {"m0": float(((m0 - y) ** 2).sum() / cells), "uniform_z": float(((zhat - y) ** 2).sum() / cells),
            "uniform_best": float(((np.broadcast_to(best, y.shape) - y) ** 2).sum() / cells)}


def train(net, data, epochs=20, batch=32, lr=1e-3, seed=0, log=None, holdout=0.1, lr_min=None, keep_best=True):
    """Adam on shuffled mini-batches; returns (train losses per epoch, held-out loss per epoch).
    With lr_min the step size follows a cosine from lr down to lr_min over the epochs.  With
    keep_best (and a held-out set) the network ends with the parameters of its best held-out epoch:
    on the C4 films the held-out loss is lowest after ~30 epochs and rises after, while the
    training loss keeps falling (the network memorises the training films)."""
    torch = _torch()
    X = torch.from_numpy(data["x"]).cuda()
    Y = torch.from_numpy(data["y"]).cuda()
    n = X.shape[0]
    hold, tr = split(n, holdout, seed)
    nh = len(hold)
    hist, hhist = [], []
    best = (math.inf, None)
    for ep in range(epochs):
        order = np.random.default_rng([seed, ep]).permutation(tr)
        lr_ep = lr if lr_min is None else lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * ep / max(epochs, 1)))
        tot, cnt = 0.0, 0
        for b in range(0, len(order) - batch + 1, batch):
            sel = torch.from_numpy(order[b:b + batch]).cuda()
            tot += net.train_step(X[sel].contiguous(), Y[sel].contiguous(), lr_ep)
            cnt += 1
        hist.append(tot / max(cnt, 1))
        hl = 0.0
        if nh:
            for b in range(0, nh, batch):
                sel = torch.from_numpy(hold[b:b + batch]).cuda()
                hl += net.gradients(X[sel].contiguous(), Y[sel].contiguous())[0] * len(sel)
            hl /= nh
        hhist.append(hl)
        if keep_best and nh and hl < best[0]:
            best = (hl, net.get_params())
        if log:
            log(f"[train] epoch {ep}: train loss {hist[-1]:.4f}, held-out {hl:.4f}")
    if best[1] is not None:
        net.set_params(best[1])
        if log:
            log(f"[train] kept the parameters of epoch {int(np.argmin(hhist))} (held-out {best[0]:.4f})")
    return hist, hhist


def texture(state):
    """|<m>| of a relaxed film: 1 for a uniform state, well below for domains, labyrinths and
    skyrmion lattices."""
    return float(np.linalg.norm(state.mean_m()))


def evaluate(net, n=4, size=256, seed=7, torque_tol=1e-6, maxiter=200000, log=None, perturb=None):
    """Plain LBFGS from a random state vs network warm start + LBFGS from the same state, both to
    max|m x H| / max|H| <= torque_tol.  Returns one record per sample.

    With perturb = eps, plain LBFGS also runs from m0 + eps * (unit Gaussian noise), renormalised:
    the control of how far the method without the network is from reaching one and the same
    minimum when its start moves by eps (E_rel_diff_perturbed)."""
    mesh = Mesh(size, size, 1, CELL, CELL, CELL)
    out = []
    for i in range(n):
        rng = np.random.default_rng([seed, i])
        p = random_params(rng)
        m0 = random_m(rng, size, size)
        st_plain, s_plain, t_plain = relax(mesh, m0, p, torque_tol, maxiter)
        st_warm, s_warm, t_warm = relax(mesh, m0, p, torque_tol, maxiter, net=net)
        rel = abs(s_warm["E"] - s_plain["E"]) / max(abs(s_plain["E"]), 1e-300)
        rec = dict(p, t_plain=t_plain, t_warm=t_warm, speedup=t_plain / t_warm, E_plain=s_plain["E"],
                   E_warm=s_warm["E"], E_rel_diff=rel, it_plain=s_plain["iterations"], it_warm=s_warm["iterations"],
                   stop_plain=s_plain["stop_reason"], stop_warm=s_warm["stop_reason"],
                   mean_m_plain=texture(st_plain), mean_m_warm=texture(st_warm))
        if perturb:
            mp = m0 + perturb * np.random.default_rng([seed, i, 1]).standard_normal(m0.shape)
            mp /= np.linalg.norm(mp, axis=-1, keepdims=True)
            _, s_pert, t_pert = relax(mesh, mp, p, torque_tol, maxiter)
            rec.update(E_perturbed=s_pert["E"], it_perturbed=s_pert["iterations"],
                       E_rel_diff_perturbed=abs(s_pert["E"] - s_plain["E"]) / max(abs(s_plain["E"]), 1e-300))
        out.append(rec)
        if log:
            extra = f", perturbed start dE/E {rec['E_rel_diff_perturbed']:.2e}" if perturb else ""
            log(f"[eval] {i}: plain {t_plain:.2f} s ({s_plain['iterations']} it, |<m>| {rec['mean_m_plain']:.3f}), "
                f"warm {t_warm:.2f} s ({s_warm['iterations']} it, |<m>| {rec['mean_m_warm']:.3f}), "
                f"speedup {rec['speedup']:.2f}, dE/E {rel:.2e}{extra}")
    return out
