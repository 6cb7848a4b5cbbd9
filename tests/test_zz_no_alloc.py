"""mafx_mean_m_noalloc (the entry point that allocates nothing) against mafx_mean_m through the
C-ABI: bit-identical <m> on two calls in a row, the second on a smaller film, so partial sums the
first call left in the shared static array must not leak into the second.  Both films hold a few
zero vectors, so the average runs over the non-zero cells only (state.cpp:32-45)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _film(shape, seed):
    m = np.random.default_rng(seed).standard_normal(tuple(shape) + (3,))
    m[0, 0, 0] = 0.0
    m[shape[0] // 2, shape[1] - 1, shape[2] - 1] = 0.0
    return m


def test_mean_m_noalloc_matches_mean_m():
    import torch

    import magnumaf as M
    from magnumaf import _abi

    L = _abi.lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda")
    for shape, seed in [((256, 64, 4), 11), ((17, 5, 1), 12)]:
        print(f"\n[mean_m_noalloc] cells={shape} seed={seed}", flush=True)
        mesh = M.Mesh(*shape, 2e-9, 2e-9, 2e-9)
        state = M.State(mesh, 1.1e6, _film(shape, seed))
        ref = np.zeros(3)
        ctx = state._context([M.ExchangeField(1.5e-11)], "f64")
        _abi.check(L.mafx_mean_m(ctx.h, C.c_void_p(state.m_dev.data_ptr()), ref.ctypes.data_as(C.POINTER(C.c_double))))
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        rc = L.mafx_mean_m_noalloc(C.c_void_p(state.m_dev.data_ptr()), mesh.n, C.c_void_p(out.data_ptr()), stream)
        assert rc == 0, L.mafx_last_error().decode()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.array_equal(got, ref), (got, ref)
        # and both against numpy over the non-zero cells
        m = state.m
        nz = np.linalg.norm(m, axis=-1) != 0
        assert np.max(np.abs(got - m[nz].mean(axis=0))) <= 1e-15
    # sizes outside [1, 2^60] are refused before anything is launched
    assert L.mafx_mean_m_noalloc(C.c_void_p(state.m_dev.data_ptr()), 0, C.c_void_p(out.data_ptr()), stream) != 0
    assert b"ncells" in L.mafx_last_error()
