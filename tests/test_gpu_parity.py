"""GPU parity: the HIP path (through the C-ABI, via the magnumaf host layer) against the FP64 CPU
oracle on the same seeded inputs.

Tolerances: the FP64 path must match H_eff normwise, max|dH| / max|H_ref| <= 1e-12; energies to
1e-12 relative to the largest term energy; trajectories step for step.  The FP32-field path's
error is measured and bounded at 1e-5 normwise."""
import math

import numpy as np
import pytest

import oracle_lib as O
from golden_values import known

pytestmark = pytest.mark.gpu

MU0 = O.MU0
FP64_TOL = 1e-12


def mf():
    import magnumaf
    return magnumaf


def rand_m(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(tuple(shape) + (3,))


def normwise(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def energy_scale(mesh, Ms, E_terms):
    """max(|E_term|, mu0 Ms^2 V): a vanishing term energy is rounding noise of its own size."""
    V = mesh[0] * mesh[1] * mesh[2] * mesh[3] * mesh[4] * mesh[5]
    return max(float(np.max(np.abs(E_terms))), MU0 * Ms * Ms * V)


def build(mesh, Ms, m, specs, Ms_field=None):
    """specs: list of (kind, kwargs) -> (gpu State, gpu terms, oracle Sim)"""
    M = mf()
    g = M.Mesh(*mesh)
    gterms, oterms = [], []
    for kind, kw in specs:
        if kind == "demag":
            gterms.append(M.DemagField(g))
            oterms.append(O.term(O.DEMAG))
        elif kind == "exchange":
            gterms.append(M.ExchangeField(kw["A"]))
            oterms.append(O.term(O.EXCHANGE, kw["A"]))
        elif kind == "dmi":
            gterms.append(M.DmiField(kw["D"], kw.get("axis", (0, 0, 1))))
            oterms.append(O.term(O.DMI, kw["D"], axis=kw.get("axis", (0, 0, 1))))
        elif kind == "uniaxial":
            gterms.append(M.UniaxialAnisotropyField(kw["K"], kw.get("axis", (0, 0, 1))))
            oterms.append(O.term(O.UNIAXIAL, kw["K"], axis=kw.get("axis", (0, 0, 1))))
        elif kind == "zeeman":
            gterms.append(M.ExternalField(kw["H"]))
            oterms.append(O.term(O.ZEEMAN, array0=O.to_flat(kw["H"])))
        elif kind == "zeeman_h":
            gterms.append(M.ExternalField(homogeneous=kw["h"]))
            oterms.append(O.term(O.ZEEMAN, field=kw["h"]))
        elif kind == "rkky":
            gterms.append(M.RKKYExchangeField(kw["rk"], kw["A"], g, kw.get("idx")))
            idx = None if kw.get("idx") is None else O.scalar_flat(kw["idx"]).astype(np.uint32)
            oterms.append(O.term(O.RKKY, array0=O.scalar_flat(kw["rk"]), array1=O.scalar_flat(kw["A"]), index0=idx))
        elif kind == "atom_dipole":
            gterms.append(M.AtomisticDipoleDipoleField(g))
            oterms.append(O.term(O.ATOM_DIPOLE))
        elif kind == "atom_exchange":
            gterms.append(M.AtomisticExchangeField(kw["J"]))
            oterms.append(O.term(O.ATOM_EXCHANGE, kw["J"]))
        elif kind == "atom_dmi":
            gterms.append(M.AtomisticDmiField(kw["D"], kw.get("axis", (0, 0, -1))))
            oterms.append(O.term(O.ATOM_DMI, kw["D"], axis=kw.get("axis", (0, 0, -1))))
        elif kind == "atom_uniaxial":
            gterms.append(M.AtomisticUniaxialAnisotropyField(kw["K"], kw.get("axis", (0, 0, 1))))
            oterms.append(O.term(O.ATOM_UNIAXIAL, kw["K"], axis=kw.get("axis", (0, 0, 1))))
        elif kind == "atom_zeeman_h":
            gterms.append(M.AtomisticExternalField(homogeneous=kw["h"]))
            oterms.append(O.term(O.ATOM_ZEEMAN, field=kw["h"]))
        else:
            raise ValueError(kind)
    state = M.State(g, Ms if Ms_field is None else Ms_field, m)
    sim = O.Sim(mesh, Ms, m, oterms, Ms_field=Ms_field)
    return state, gterms, sim


MICRO_ALL = [("demag", {}), ("exchange", {"A": 1.5e-11}), ("dmi", {"D": 3e-3}),
             ("uniaxial", {"K": 8e5, "axis": (0.1, 0.2, 1.0)})]


@pytest.mark.parametrize("mesh", [(5, 4, 3, 1e-9, 1.2e-9, 0.8e-9), (7, 3, 1, 2e-9, 1e-9, 3e-9),
                                  (16, 8, 4, 2e-9, 2e-9, 2e-9), (100, 25, 1, 5e-9, 5e-9, 3e-9),
                                  (33, 7, 5, 1e-9, 2e-9, 1.5e-9), (1, 1, 1, 2e-9, 2e-9, 2e-9)])
@pytest.mark.parametrize("which", ["demag", "exchange", "dmi", "uniaxial", "all"])
def test_heff_fp64_matches_oracle(mesh, which):
    print(f"\n[heff fp64] mesh={mesh} term={which}", flush=True)
    specs = MICRO_ALL if which == "all" else [s for s in MICRO_ALL if s[0] == which]
    m = rand_m(mesh[:3], 7)
    state, terms, sim = build(mesh, 1.1e6, m, specs)
    M = mf()
    llg = M.LLGIntegrator(0.1, terms)
    h = llg.H_in_Apm(state)
    ref = sim.heff()
    assert normwise(h, ref) <= FP64_TOL, normwise(h, ref)
    E_gpu = llg.Eeff_in_J(state)
    E_ref, Et_ref = sim.energy()
    assert abs(E_gpu - E_ref) <= FP64_TOL * energy_scale(mesh, 1.1e6, Et_ref)


@pytest.mark.parametrize("mesh", [(8, 256, 2, 2e-9, 1e-9, 2e-9), (5, 512, 3, 1e-9, 1e-9, 1e-9),
                                  (4, 1024, 1, 2e-9, 2e-9, 3e-9), (3, 1024, 4, 2e-9, 2e-9, 2e-9),
                                  (256, 8, 2, 1e-9, 2e-9, 2e-9), (512, 4, 1, 2e-9, 2e-9, 3e-9),
                                  (1024, 2, 3, 2e-9, 1e-9, 2e-9), (2048, 1, 1, 2e-9, 2e-9, 2e-9),
                                  (256, 256, 1, 3e-9, 3e-9, 3e-9)])
def test_heff_fp64_pow2_axes(mesh):
    """Power-of-two nx of 256..2048 or ny of 256..1024: the x or y transforms run the
    compile-time shaped kernels (row groups of one plane, whole column tiles).  The energy on rows
    of whole work-groups (nx >= 256) is summed on the z-column lanes with per-group partials."""
    print(f"\n[heff fp64 pow2 axes] mesh={mesh}", flush=True)
    m = rand_m(mesh[:3], 21)
    state, terms, sim = build(mesh, 1.1e6, m, MICRO_ALL)
    M = mf()
    llg = M.LLGIntegrator(0.1, terms)
    h = llg.H_in_Apm(state)
    assert normwise(h, sim.heff()) <= FP64_TOL, normwise(h, sim.heff())
    E_ref, Et_ref = sim.energy()
    assert abs(llg.Eeff_in_J(state) - E_ref) <= FP64_TOL * energy_scale(mesh, 1.1e6, Et_ref)
    for i, t in enumerate(terms):
        assert abs(t.Energy_in_J(state) - Et_ref[i]) <= FP64_TOL * energy_scale(mesh, 1.1e6, Et_ref), i


@pytest.mark.parametrize("precision,tol", [("f64", FP64_TOL), ("f32", 1e-5)])
def test_ms_field_airbox_pow2_x(precision, tol):
    """Ms_field scaling of the packed x input on the compile-time shaped x kernels (nx = 256)."""
    nx, ny, nz = 256, 16, 2
    m = np.zeros((nx, ny, nz, 3))
    m[64:192, 4:12, :, 0] = 1
    m[64:192, 4:12, :, 1] = 0.5
    mesh = (nx, ny, nz, 2e-9, 2e-9, 2e-9)
    print(f"\n[Ms_field pow2 x] mesh={mesh} {precision}", flush=True)
    M = mf()
    g = M.Mesh(*mesh)
    state = M.State(g, 1e5, m)
    assert state.Ms_field is not None
    sim = O.Sim(mesh, 1e5, m, [O.term(O.DEMAG), O.term(O.EXCHANGE, 1e-11)])
    llg = M.LLGIntegrator(0.1, [M.DemagField(g), M.ExchangeField(1e-11)], precision=precision)
    assert normwise(llg.H_in_Apm(state), sim.heff()) <= tol


def test_heff_each_term_energy_and_zeeman():
    mesh = (12, 10, 4, 2e-9, 2e-9, 2e-9)
    nx, ny, nz = mesh[:3]
    rng = np.random.default_rng(3)
    H = rng.standard_normal((nx, ny, nz, 3)) * 1e5
    specs = MICRO_ALL + [("zeeman", {"H": H})]
    m = rand_m(mesh[:3], 9)
    state, terms, sim = build(mesh, 8e5, m, specs)
    _, Et_ref = sim.energy()
    for i, t in enumerate(terms):
        h = t.H_in_Apm(state)
        ref = sim.heff(mask=1 << i)
        assert normwise(h, ref) <= FP64_TOL, (i, normwise(h, ref))
        e = t.Energy_in_J(state)
        assert abs(e - Et_ref[i]) <= FP64_TOL * energy_scale(mesh, 8e5, Et_ref[i]), (i, e, Et_ref[i])


def test_ms_field_airbox():
    # zero-|m| cells build Ms_field (state.cpp:71-89): demag cube energy in an air box
    x = y = z = 1e-9
    nx, ny, nz = 12, 10, 8
    m = np.zeros((2 * nx, 2 * ny, 2 * nz, 3))
    m[nx // 2:2 * nx - nx // 2, ny // 2:2 * ny - ny // 2, nz // 2:2 * nz - nz // 2, 0] = 1
    mesh = (2 * nx, 2 * ny, 2 * nz, x / nx, y / ny, z / nz)
    M = mf()
    g = M.Mesh(*mesh)
    state = M.State(g, 1e5, m)
    assert state.Ms_field is not None
    demag = M.DemagField(g)
    E = demag.Energy_in_J(state)
    analytic = 1.0 / 6.0 * x * y * z * 1e5**2 * MU0
    assert abs(E - analytic) <= (E + analytic) / 2 * 5.3e-8  # magnumaf/test/micro_demag.cpp:31-55
    sim = O.Sim(mesh, 1e5, m, [O.term(O.DEMAG), O.term(O.EXCHANGE, 1e-11)])
    ex = M.ExchangeField(1e-11)
    llg = M.LLGIntegrator(0.1, [demag, ex])
    assert normwise(llg.H_in_Apm(state), sim.heff()) <= FP64_TOL
    assert abs(llg.Eeff_in_J(state) - sim.energy()[0]) <= FP64_TOL * abs(sim.energy()[1]).max()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_demag_tensor_cache(precision, tmp_path, monkeypatch):
    """DemagField(mesh, caching=True): the first context computes the tensor and writes it to the
    cache folder, the second reads it and gives the same field to the last bit; a damaged file is
    ignored (the tensor is computed again, demag_field.cpp:185-204 / cache_manager.cpp:76-95); a
    folder above its size limit is shrunk, oldest files first (cache_manager.cpp:29-48,104-110)."""
    import os
    monkeypatch.setenv("MAFX_CACHE_DIR", str(tmp_path))
    monkeypatch.delenv("MAFX_CACHE_MAX_BYTES", raising=False)
    M = mf()
    mesh = (24, 16, 2, 2e-9, 2e-9, 2e-9)
    g = M.Mesh(*mesh)
    state = M.State(g, 1e6, rand_m(mesh[:3], 31))
    sim = O.Sim(mesh, 1e6, rand_m(mesh[:3], 31), [O.term(O.DEMAG)])
    tol = FP64_TOL if precision == "f64" else 1e-5

    def field():
        d = M.DemagField(g, caching=True)
        ctx = state._context([d], precision)
        return ctx.tensor_source(), ctx.heff(state.m_dev).cpu().numpy().copy()

    src1, h1 = field()
    files = sorted(os.listdir(tmp_path))
    assert src1 == 1 and len(files) == 1
    src2, h2 = field()
    assert src2 == 2 and np.array_equal(h1, h2)
    assert normwise(h2, sim.heff_flat()) <= tol
    # flip one payload byte: the checksum rejects the file and the tensor is computed again
    path = tmp_path / files[0]
    raw = bytearray(path.read_bytes())
    raw[-5] ^= 0xFF
    path.write_bytes(bytes(raw))
    src3, h3 = field()
    assert src3 == 1 and np.array_equal(h1, h3)
    # a limit below one file: the write is followed by a shrink that removes it
    for f in os.listdir(tmp_path):
        os.remove(tmp_path / f)
    monkeypatch.setenv("MAFX_CACHE_MAX_BYTES", "1000")
    src4, _ = field()
    assert src4 == 1 and os.listdir(tmp_path) == []


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_heff_is_deterministic(precision):
    """Every kernel of the evaluation has a fixed schedule (no atomics, fixed reduction trees):
    repeated evaluations, in one context and in a fresh one, agree to the last bit, as do the
    energies and an LLG right-hand side."""
    M = mf()
    mesh = (1024, 8, 4, 2e-9, 2e-9, 2e-9)
    g = M.Mesh(*mesh)
    m = rand_m(mesh[:3], 41)
    state = M.State(g, 1.1e6, m)
    terms = [M.DemagField(g), M.ExchangeField(1.5e-11), M.DmiField(3e-3), M.UniaxialAnisotropyField(8e5)]
    ctx = state._context(terms, precision)
    h1 = ctx.heff(state.m_dev).cpu().numpy().copy()
    h2 = ctx.heff(state.m_dev).cpu().numpy().copy()
    e1 = ctx.energy(state.m_dev)
    state2 = M.State(g, 1.1e6, m)
    ctx2 = state2._context(list(terms), precision)
    h3 = ctx2.heff(state2.m_dev).cpu().numpy().copy()
    e2 = ctx2.energy(state2.m_dev)
    assert np.array_equal(h1.view(np.uint64), h2.view(np.uint64))
    assert np.array_equal(h1.view(np.uint64), h3.view(np.uint64))
    assert e1[0] == e2[0] and np.array_equal(e1[1], e2[1])


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("mesh,ms_field", [((256, 256, 2, 2e-9, 2e-9, 2e-9), False),
                                           ((1024, 64, 2, 2e-9, 1e-9, 2e-9), False),
                                           ((512, 512, 1, 1e-9, 1e-9, 2e-9), True)])
def test_persistent_passes_bitwise(mesh, ms_field, precision, monkeypatch):
    """The x forward pass (and the FP32 y forward pass) run one resident set of groups sized by
    the CU count, each taking tiles g, g + grid, ... and prefetching the next.  With the CU count
    forced to 1 every group takes many tiles; the field must equal the default launch (one tile
    per group on these meshes) to the last bit, also with an Ms field (the x forward's scaled
    input variant)."""
    M = mf()
    print(f"\n[persistent passes] mesh={mesh} ms_field={ms_field} {precision}", flush=True)
    g = M.Mesh(*mesh)
    m = rand_m(mesh[:3], 7)
    if ms_field:
        m[: mesh[0] // 4] = 0  # an air region: State derives an Ms field from |m| = 0 cells
    terms = lambda: [M.DemagField(g), M.ExchangeField(1.5e-11), M.DmiField(3e-3)]
    state = M.State(g, 1.1e6, m)
    assert (state.Ms_field is not None) == ms_field
    h_default = state._context(terms(), precision).heff(state.m_dev).cpu().numpy().copy()
    monkeypatch.setenv("MAFX_NUM_CUS", "1")
    state1 = M.State(g, 1.1e6, m)
    h_one = state1._context(terms(), precision).heff(state1.m_dev).cpu().numpy().copy()
    assert np.array_equal(h_default.view(np.uint64), h_one.view(np.uint64))


def test_demag_cube_energy_gpu():
    # magnumaf/test/micro_demag.cpp:6-29
    x = y = z = 1e-9
    nx, ny, nz = 8, 10, 12
    M = mf()
    g = M.Mesh(nx, ny, nz, x / nx, y / ny, z / nz)
    m = np.zeros((nx, ny, nz, 3))
    m[..., 0] = 1
    E = M.DemagField(g).Energy_in_J(M.State(g, 1e5, m))
    analytic = 1.0 / 6.0 * x * y * z * 1e5**2 * MU0
    assert abs(E - analytic) <= (E + analytic) / 2 * 5.3e-8
    # the FP32 path (tensor built on the GPU, symmetric real spectrum, FP32 transforms)
    E32 = M.DemagField(g).Energy_in_J(M.State(g, 1e5, m), precision="f32")
    assert abs(E32 - analytic) <= analytic * 1e-5, (E32, analytic)


# FP32 field path: nz = 1 (k-space product), 2 <= nz <= 8 (fused z stage), nz > 8 (strided z
# transforms), power-of-two y lines from 1 to 1024 and odd and prime sizes.  Tolerance: FP32
# rounding of the FFTs and stencils.
FP32_TOL = 1e-5
FP32_TOL_LONG_ROWS = 3e-5


@pytest.mark.parametrize("mesh", [(16, 8, 4, 2e-9, 2e-9, 2e-9), (64, 32, 1, 3e-9, 3e-9, 3e-9),
                                  (5, 4, 3, 1e-9, 1.2e-9, 0.8e-9), (24, 16, 2, 1e-9, 2e-9, 1e-9),
                                  (16, 8, 5, 2e-9, 1e-9, 1e-9), (128, 256, 4, 2e-9, 2e-9, 2e-9),
                                  (8, 1024, 4, 2e-9, 2e-9, 2e-9), (10, 64, 6, 1e-9, 1e-9, 2e-9),
                                  (7, 3, 1, 2e-9, 1e-9, 3e-9), (33, 7, 5, 1e-9, 2e-9, 1.5e-9),
                                  (12, 6, 10, 2e-9, 2e-9, 2e-9), (20, 9, 8, 1e-9, 1e-9, 1e-9),
                                  (6, 512, 2, 2e-9, 1e-9, 2e-9), (12, 256, 1, 2e-9, 2e-9, 3e-9),
                                  (256, 16, 2, 1e-9, 2e-9, 2e-9), (512, 8, 1, 2e-9, 2e-9, 3e-9),
                                  (1024, 4, 2, 2e-9, 1e-9, 2e-9), (2048, 2, 1, 2e-9, 2e-9, 2e-9),
                                  (1, 1, 1, 2e-9, 2e-9, 2e-9), (16, 16, 7, 2e-9, 2e-9, 2e-9),
                                  (64, 32, 8, 2e-9, 2e-9, 2e-9)])
def test_heff_fp32_field_error(mesh):
    print(f"\n[heff fp32] mesh={mesh}", flush=True)
    m = rand_m(mesh[:3], 5)
    state, terms, sim = build(mesh, 1.1e6, m, MICRO_ALL)
    M = mf()
    h32 = M.LLGIntegrator(0.1, terms, precision="f32").H_in_Apm(state)
    err = normwise(h32, sim.heff())
    print(f"[fp32 field] mesh={mesh} normwise error={err:.3e}", flush=True)
    assert err <= FP32_TOL
    # the demag term alone, where the symmetric tensor spectrum acts.  Rows of 2048 cells (4096
    # padded points) strips carry more FP32 rounding relative to their small demag field: 1.4e-5
    # at 2048x2x1 and 2.5e-5 at 2048x4x1, the same to the last bit on the generic (run-time shaped)
    # and the compile-time shaped x kernels
    hd = M.LLGIntegrator(0.1, [terms[0]], precision="f32").H_in_Apm(state)
    err_d = normwise(hd, sim.heff(mask=1))
    assert err_d <= (FP32_TOL if mesh[0] < 2048 else FP32_TOL_LONG_ROWS), err_d


@pytest.mark.parametrize("mesh", [(20, 9, 8, 1e-9, 1e-9, 1e-9), (64, 32, 8, 2e-9, 2e-9, 2e-9),
                                  (16, 16, 7, 2e-9, 2e-9, 2e-9), (16, 8, 5, 2e-9, 1e-9, 1e-9),
                                  (24, 16, 2, 1e-9, 2e-9, 1e-9), (64, 32, 1, 3e-9, 3e-9, 3e-9),
                                  (12, 6, 10, 2e-9, 2e-9, 2e-9)])
def test_fp32_full_tensor(mesh):
    """DemagField(full_tensor=True): the FP32 field on the reference tensor's spectrum (symmetric
    quarter + FP16 residual) against the oracle, on the fused z stages (the direct DFT at nz = 2, 5
    and 7, the in-place FFT form at nz = 8), the nz = 1 product and the strided z passes (nz = 10)."""
    print(f"\n[fp32 full tensor] mesh={mesh}", flush=True)
    m = rand_m(mesh[:3], 8)
    state, terms, sim = build(mesh, 1.1e6, m, MICRO_ALL)
    M = mf()
    g = state.mesh
    hd = M.LLGIntegrator(0.1, [M.DemagField(g, full_tensor=True)], precision="f32").H_in_Apm(state)
    err = normwise(hd, sim.heff(mask=1))
    print(f"[fp32 full tensor] mesh={mesh} demag normwise error={err:.3e}", flush=True)
    assert err <= FP32_TOL


def test_fp32_host_tensor():
    """DemagField(host_tensor=True) on the FP32-field path: the real-space tensor comes from the host
    in the reference's operation order (the FP64 path's), so the field matches the oracle within the
    FP32 tolerance and differs from the default device-built tensor only by far-field rounding."""
    M = mf()
    mesh = (256, 16, 2, 2e-9, 2e-9, 2e-9)
    print(f"\n[fp32 host tensor] mesh={mesh}", flush=True)
    m = rand_m(mesh[:3], 19)
    g = M.Mesh(*mesh)
    sim = O.Sim(mesh, 8e5, m, [O.term(O.DEMAG)])
    href = sim.heff()
    h_host = M.LLGIntegrator(0.1, [M.DemagField(g, host_tensor=True)], precision="f32").H_in_Apm(M.State(g, 8e5, m))
    h_dev = M.LLGIntegrator(0.1, [M.DemagField(g)], precision="f32").H_in_Apm(M.State(g, 8e5, m))
    assert normwise(h_host, href) <= FP32_TOL, normwise(h_host, href)
    assert normwise(h_host, h_dev) <= FP32_TOL, normwise(h_host, h_dev)


def test_fp32_ms_field_airbox():
    x = y = z = 1e-9
    nx, ny, nz = 6, 5, 4
    m = np.zeros((2 * nx, 2 * ny, 2 * nz, 3))
    m[nx // 2:2 * nx - nx // 2, ny // 2:2 * ny - ny // 2, nz // 2:2 * nz - nz // 2, 0] = 1
    mesh = (2 * nx, 2 * ny, 2 * nz, x / nx, y / ny, z / nz)
    print(f"\n[fp32 Ms_field] mesh={mesh}", flush=True)
    M = mf()
    g = M.Mesh(*mesh)
    state = M.State(g, 1e5, m)
    sim = O.Sim(mesh, 1e5, m, [O.term(O.DEMAG), O.term(O.EXCHANGE, 1e-11)])
    llg = M.LLGIntegrator(0.1, [M.DemagField(g), M.ExchangeField(1e-11)], precision="f32")
    assert normwise(llg.H_in_Apm(state), sim.heff()) <= FP32_TOL


@pytest.mark.parametrize("mesh", [(16, 12, 1, 2.7e-10, 2.7e-10, 2.7e-10), (8, 6, 3, 2.7e-10, 2.7e-10, 2.7e-10),
                                  (12, 16, 2, 2.7e-10, 2.7e-10, 2.7e-10)])
def test_fp32_atomistic_dipole(mesh):
    print(f"\n[fp32 atomistic] mesh={mesh}", flush=True)
    J = 1.6e-21
    specs = [("atom_dipole", {}), ("atom_exchange", {"J": J})]
    m = rand_m(mesh[:3], 17)
    state, terms, sim = build(mesh, 2.6 * 9.274009994e-24, m, specs)
    M = mf()
    llg = M.LLGIntegrator(0.1, terms, precision="f32")
    assert normwise(llg.H_in_Apm(state), sim.heff()) <= FP32_TOL
    hd = M.LLGIntegrator(0.1, [terms[0]], precision="f32").H_in_Apm(state)
    assert normwise(hd, sim.heff(mask=1)) <= FP32_TOL


def test_rkky_energy_sweep_gpu():
    # magnumaf/test/rkky.cpp:11-55 (mumax3 RKKY example) on the GPU path
    nx, ny, nz, dx = 10, 10, 2, 1e-9
    mesh = (nx, ny, nz, dx, dx, dx)
    rk = np.full((nx, ny, nz), -1e-3 * dx / 2.0)
    A = np.full((nx, ny, nz), 10e-12)
    m = np.zeros((nx, ny, nz, 3))
    m[..., 0] = 1
    state, terms, sim = build(mesh, 1e6, m, [("demag", {}), ("rkky", {"rk": rk, "A": A})])
    M = mf()
    llg = M.LLGIntegrator(1.0, terms)
    Es, Eo = [], []
    for i in range(0, 360, 3):
        mm = m.copy()
        mm[:, :, 1, 0] = math.cos(i * math.pi / 180.0)
        mm[:, :, 1, 1] = math.sin(i * math.pi / 180.0)
        state.set_m_raw(mm)
        sim.set_m(mm, normalize=False)
        Es.append(llg.Eeff_in_J(state))
        Eo.append(sim.energy()[0])
    Es, Eo = np.array(Es), np.array(Eo)
    assert np.max(np.abs(Es - Eo)) <= FP64_TOL * np.max(np.abs(Eo))
    h = llg.H_in_Apm(state)
    assert normwise(h, sim.heff()) <= FP64_TOL


def test_rkky_layers_with_indices():
    nx, ny, nz = 6, 5, 4
    mesh = (nx, ny, nz, 2e-9, 2e-9, 2e-9)
    rng = np.random.default_rng(1)
    rk = np.zeros((nx, ny, nz))
    rk[:, :, 1:3] = -1e-3 * 2e-9 / 2
    A = 1e-11 * (1 + 0.5 * rng.random((nx, ny, nz)))
    idx = np.zeros((nx, ny, nz), dtype=np.uint32)
    idx[:, :, 2:] = 1
    m = rand_m((nx, ny, nz), 4)
    state, terms, sim = build(mesh, 1e6, m, [("rkky", {"rk": rk, "A": A, "idx": idx})])
    assert normwise(terms[0].H_in_Apm(state), sim.heff()) <= FP64_TOL


@pytest.mark.parametrize("mesh", [(6, 5, 8, 2e-9, 2e-9, 1e-9), (256, 2, 8, 2e-9, 3e-9, 2e-9)])
def test_rkky_precomputed_weights_layers(mesh):
    """The RKKY neighbour weights are precomputed per term (FP64 expressions, stored in the
    stencils' type): eight layers with layer indices, RKKY entries on some layers only, a
    non-uniform exchange constant with zeros, on the generic per-cell kernel (nx = 6) and the
    z-column one (nx = 256).  FP64 weights hold the 1e-12 bound; the FP32 rounding of each weight
    is within the FP32 tolerance.  Then new arrays assigned to the same term object: the cached
    context uploads them (mafx_ctx_update_term) and its weights follow."""
    nx, ny, nz = mesh[:3]
    print(f"\n[rkky weights] mesh={mesh}", flush=True)
    rng = np.random.default_rng(21)
    rk = np.zeros((nx, ny, nz))
    rk[:, :, 1:6] = -1e-3 * mesh[5] / 2 * (1 + rng.random((nx, ny, 5)))
    A = 1e-11 * (1 + rng.random((nx, ny, nz)))
    A[0, :, 3] = 0.0
    idx = np.zeros((nx, ny, nz), dtype=np.uint32)
    idx[:, :, 3:] = 1
    idx[:, :, 6:] = 2
    m = rand_m((nx, ny, nz), 22)
    M = mf()
    for prec, tol in [("f64", FP64_TOL), ("f32", FP32_TOL)]:
        state, terms, sim = build(mesh, 1e6, m, [("rkky", {"rk": rk, "A": A, "idx": idx})])
        h = M.LLGIntegrator(0.05, terms, precision=prec).H_in_Apm(state)
        assert normwise(h, sim.heff()) <= tol, (prec, normwise(h, sim.heff()))
    llg = M.LLGIntegrator(0.05, terms, precision="f32")
    rk2 = np.full((nx, ny, nz), -2e-3 * mesh[5] / 2)
    A2 = np.full((nx, ny, nz), 1.3e-11)
    g = M.Mesh(*mesh)
    terms[0].rk = M._scalar_flat(rk2, g)
    terms[0].A = M._scalar_flat(A2, g)
    terms[0].idx = None
    h = llg.H_in_Apm(state)
    _, _, sim2 = build(mesh, 1e6, m, [("rkky", {"rk": rk2, "A": A2})])
    assert normwise(h, sim2.heff()) <= FP32_TOL, normwise(h, sim2.heff())
    assert normwise(h, sim.heff()) > 1e-2  # the old arrays' field differs


@pytest.mark.parametrize("precision,tol", [("f64", FP64_TOL), ("f32", 1e-5)])
def test_rkky_multilayer_rows(precision, tol):
    """The C5 term list (demag, RKKY between layers, DMI, uniaxial, Zeeman) and RKKY alone on
    rows of 256 cells: the z-column per-cell kernel with a compile-time and a run-time term list."""
    nx, ny, nz = 256, 3, 4
    mesh = (nx, ny, nz, 2e-9, 2e-9, 2e-9)
    print(f"\n[rkky rows] mesh={mesh} {precision}", flush=True)
    rng = np.random.default_rng(8)
    rk = np.full((nx, ny, nz), -1e-3 * 2e-9 / 2)
    A = np.full((nx, ny, nz), 1.3e-11)
    m = rand_m((nx, ny, nz), 14)
    specs = [("demag", {}), ("rkky", {"rk": rk, "A": A}), ("dmi", {"D": 2.5e-3}),
             ("uniaxial", {"K": 7e5, "axis": (0.0, 0.0, 1.0)}), ("zeeman_h", {"h": (50e-3 / MU0, 0, 0)})]
    state, terms, sim = build(mesh, 1e6, m, specs)
    M = mf()
    h = M.LLGIntegrator(0.05, terms, precision=precision).H_in_Apm(state)
    assert normwise(h, sim.heff()) <= tol, normwise(h, sim.heff())
    A2 = 1e-11 * (1 + 0.5 * rng.random((nx, ny, nz)))
    idx = np.zeros((nx, ny, nz), dtype=np.uint32)
    idx[:, :, 2:] = 1
    state, terms, sim = build(mesh, 1e6, m, [("rkky", {"rk": rk, "A": A2, "idx": idx})])
    h = M.LLGIntegrator(0.05, terms, precision=precision).H_in_Apm(state)
    assert normwise(h, sim.heff()) <= tol, normwise(h, sim.heff())


@pytest.mark.parametrize("mesh", [(2, 1, 1, 2.7e-10, 2.7e-10, 2.7e-10), (16, 12, 1, 2.7e-10, 2.7e-10, 2.7e-10),
                                  (8, 6, 3, 2.7e-10, 2.7e-10, 2.7e-10), (256, 3, 2, 2.7e-10, 2.7e-10, 2.7e-10)])
def test_atomistic_terms(mesh):
    J = 1.6e-21
    specs = [("atom_dipole", {}), ("atom_exchange", {"J": J}), ("atom_dmi", {"D": 0.2 * J}),
             ("atom_uniaxial", {"K": 0.02 * J}), ("atom_zeeman_h", {"h": (0, 0, 1 / MU0)})]
    mus = 2.6 * 9.274009994e-24
    m = rand_m(mesh[:3], 13)
    state, terms, sim = build(mesh, mus, m, specs)
    M = mf()
    llg = M.LLGIntegrator(0.1, terms)
    assert normwise(llg.H_in_Apm(state), sim.heff()) <= FP64_TOL
    E_ref, Et = sim.energy()
    assert abs(llg.Eeff_in_J(state) - E_ref) <= FP64_TOL * np.max(np.abs(Et))


def test_vector_helpers():
    import ctypes as C
    import torch
    M = mf()
    from magnumaf._abi import check, lib
    g = M.Mesh(37, 11, 3, 1e-9, 1e-9, 1e-9)
    rng = np.random.default_rng(2)
    a = rng.standard_normal(3 * g.n)
    b = rng.standard_normal(3 * g.n)
    state = M.State(g, 8e5, rand_m((37, 11, 3), 2))
    ctx = state._util_ctx()
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = C.c_double()
    check(lib().mafx_dot(ctx.h, C.c_void_p(da.data_ptr()), C.c_void_p(db.data_ptr()), C.byref(out)))
    assert abs(out.value - float(np.dot(a, b))) <= 1e-13 * np.sum(np.abs(a * b))
    check(lib().mafx_max_abs(ctx.h, C.c_void_p(da.data_ptr()), C.byref(out)))
    assert out.value == np.max(np.abs(a))
    # LLG right-hand side against the oracle's restatement of equations.hpp
    h = torch.from_numpy(b * 1e5).cuda()
    k = torch.empty_like(h)
    check(lib().mafx_llg_rhs(ctx.h, 0.3, 1e-13, C.c_void_p(state.m_dev.data_ptr()), C.c_void_p(h.data_ptr()),
                             C.c_void_p(k.data_ptr())))
    mm = state.m_dev.cpu().numpy().reshape(3, -1)
    hh = b.reshape(3, -1) * 1e5
    mxh = np.cross(mm.T, hh.T).T
    mxmxh = np.cross(mm.T, mxh.T).T
    gam = 1.760859644e11 * MU0
    kref = 1e-13 * (-gam / (1 + 0.09) * mxh - 0.3 * gam / (1 + 0.09) * mxmxh)
    assert np.max(np.abs(k.cpu().numpy().reshape(3, -1) - kref)) <= 1e-13 * np.max(np.abs(kref))
    mean = state.mean_m()
    assert np.allclose(mean, mm.mean(axis=1), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("scheme,mesh", [("DP45", (32, 8, 1, 4e-9, 4e-9, 3e-9)),
                                         ("RKF45", (32, 8, 1, 4e-9, 4e-9, 3e-9)),
                                         ("BS45", (32, 8, 1, 4e-9, 4e-9, 3e-9)),
                                         ("DP78", (32, 8, 1, 4e-9, 4e-9, 3e-9)),
                                         ("BS23", (32, 8, 1, 4e-9, 4e-9, 3e-9)),
                                         ("DP45", (512, 2, 1, 4e-9, 4e-9, 3e-9))])
def test_llg_step_replay(scheme, mesh):
    # identical controller decisions and states, step by step, on a C1-like film (the 512-cell
    # rows run the fused inverse-x / per-cell kernel)
    print(f"\n[llg replay] {scheme} mesh={mesh}", flush=True)
    nx, ny, nz = mesh[:3]
    rng = np.random.default_rng(42)
    m = np.zeros((nx, ny, nz, 3))
    m[..., 0] = 1
    m += 0.1 * rng.standard_normal(m.shape)
    h = 25e-3 / MU0
    H = (h * math.cos(math.radians(170)), h * math.sin(math.radians(170)), 0.0)
    specs = [("demag", {}), ("exchange", {"A": 1.3e-11}), ("zeeman_h", {"h": H})]
    state, terms, sim = build(mesh, 8e5, m, specs)
    M = mf()
    llg = M.LLGIntegrator(0.02, terms, scheme)
    sim.llg_init(0.02, scheme)
    # From h = 1.01e-15 the embedded error estimate of the first steps is rounding noise of the
    # stage sums, and the PI controller turns its relative differences into err^-0.14 changes of
    # h (~1e-8 between two FP64 summation orders).  The replay therefore requires the same
    # accept/reject decisions, h within 1e-6 and the state within 1e-8.  DP78's error sum has
    # weights up to +-2.5 over 13 stages and amplifies that noise: h within 1e-5 there (2e-6 is
    # seen at h ~ 3e-13), and after 40 steps the two runs sit ~1e-16 s apart in time, which at
    # |dm/dt| ~ 1e10 / s moves the state by ~3e-7, so the state is compared within 1e-6.
    htol = 1e-5 if scheme == "DP78" else 1e-6
    mtol = 1e-6 if scheme == "DP78" else 1e-8
    for i in range(40):
        r = sim.step()
        llg.step(state)
        st = llg.stats()
        assert abs(st["h_last"] - r["h"]) <= htol * r["h"], (i, st["h_last"], r["h"])
        assert abs(state.t - sim.t) <= htol * sim.t
        assert st["rejected"] == r["rejected"] and st["accepted"] == r["accepted"]
    assert np.max(np.abs(state.m - sim.m())) <= mtol


def _film_steps(mesh, precision, scheme, nsteps, seed=7):
    """nsteps steps of a C2-like film (demag, exchange, DMI, uniaxial, alpha = 0.1) from a
    near-uniform state; returns the initial and final m, the step sizes and the stats."""
    M = mf()
    nx, ny, nz = mesh[:3]
    rng = np.random.default_rng(seed)
    m = np.zeros((nx, ny, nz, 3))
    m[..., 0] = 1
    m += 0.1 * rng.standard_normal(m.shape)
    g = M.Mesh(*mesh)
    terms = [M.DemagField(g), M.ExchangeField(1.5e-11), M.DmiField(3e-3), M.UniaxialAnisotropyField(8e5)]
    state = M.State(g, 1.1e6, m)
    m0 = state.m.copy()
    llg = M.LLGIntegrator(0.1, terms, scheme, precision=precision)
    hs = []
    for _ in range(nsteps):
        llg.step(state)
        hs.append(llg.stats()["h_last"])
    return m0, state.m.copy(), np.array(hs), llg.stats(), state.t


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("scheme", ["DP45", "RKF45", "BS23"])
@pytest.mark.parametrize("mesh", [(256, 32, 4, 2e-9, 2e-9, 2e-9), (64, 16, 2, 2e-9, 2e-9, 2e-9)])
def test_rk_fused_stage_inputs_bitwise(precision, scheme, mesh, monkeypatch):
    """Stage inputs m + sum_j a_sj k_j formed in the per-cell pass that produces k_(s-1) (meshes
    from 2^20 cells) and in a pass of their own (adaptive_runge_kutta.cpp:138-164) give the same
    steps to the last bit.  MAFX_RK_FUSE_MIN_CELLS moves the threshold, so both forms run here on
    the z-column kernel's 256-cell rows and on the generic per-cell kernel."""
    print(f"\n[rk fused] {precision} {scheme} mesh={mesh}", flush=True)
    runs = []
    for thr in ["0", str(1 << 40)]:
        monkeypatch.setenv("MAFX_RK_FUSE_MIN_CELLS", thr)
        runs.append(_film_steps(mesh, precision, scheme, 12))
    (_, ma, ha, sa, ta), (_, mb, hb, sb, tb) = runs
    assert ha.tobytes() == hb.tobytes() and ta == tb
    assert sa["accepted"] == sb["accepted"] and sa["rejected"] == sb["rejected"]
    assert ma.tobytes() == mb.tobytes()


def test_fp32_stage_derivatives(monkeypatch):
    """The FP32-field integrator stores its stage derivatives k_j = dt f(m_j) in FP32 and forms
    every sum over them in FP64.  Against the same path with FP64 k_j (MAFX_RK_K64=1):

    - one DP45 step moves the state by at most 4 FP32 roundings of the step's own increment (the
      |b_j| of DP45 sum to 1.64);
    - from h = 1.01e-15 the embedded error estimate is rounding noise of the stage sums until the
      steps grow, so the controller's step sizes, and the times the steps land on, follow that
      noise in both forms; compared at the same times (integrate_dense, llg_integrator.cpp:68-104)
      the trajectories of <m> agree to the bound below, with the step counts within 5%."""
    mesh = (256, 32, 4, 2e-9, 2e-9, 2e-9)
    M = mf()
    res = {}
    for k64 in ["0", "1"]:
        monkeypatch.setenv("MAFX_RK_K64", k64)
        one = _film_steps(mesh, "f32", "DP45", 1)
        nx, ny, nz = mesh[:3]
        rng = np.random.default_rng(7)
        m = np.zeros((nx, ny, nz, 3))
        m[..., 0] = 1
        m += 0.1 * rng.standard_normal(m.shape)
        g = M.Mesh(*mesh)
        terms = [M.DemagField(g), M.ExchangeField(1.5e-11), M.DmiField(3e-3), M.UniaxialAnisotropyField(8e5)]
        state = M.State(g, 1.1e6, m)
        llg = M.LLGIntegrator(0.1, terms, "DP45", precision="f32")
        rows = llg.integrate_dense(state, 20e-12, 1e-12)
        res[k64] = (one, rows, llg.stats())
    (m0, m1_32, _, _, _), rows32, s32 = res["0"]
    (_, m1_64, _, _, _), rows64, s64 = res["1"]
    inc = np.max(np.abs(m1_64 - m0))
    d1 = np.max(np.abs(m1_32 - m1_64))
    dev = np.max(np.abs(rows32[:, 1:] - rows64[:, 1:]))
    print(f"\n[k32] one step: increment {inc:.3e}, k32 - k64 {d1:.3e} ({d1 / (inc * 2.0 ** -24):.2f} ulps); "
          f"20 ps: <m> deviation {dev:.3e}, steps {s32['accepted']} vs {s64['accepted']}, "
          f"rejected {s32['rejected']} vs {s64['rejected']}", flush=True)
    assert 0 < d1 <= 4 * 2.0 ** -24 * inc
    assert np.array_equal(rows32[:, 0], rows64[:, 0]) and len(rows32) == 21
    assert abs(s32["accepted"] - s64["accepted"]) <= 0.05 * s64["accepted"]
    assert dev <= 1e-5


@pytest.mark.parametrize("mesh", [(15, 5, 3, 2e-9, 2e-9, 2e-9), (256, 32, 4, 2e-9, 2e-9, 2e-9)])
def test_fp32_stage_derivatives_one_step(mesh, monkeypatch):
    """One DP45 step of the FP32-field path with FP32 stage derivatives against FP64 ones
    (MAFX_RK_K64=1), on an odd number of vector elements (3 x 225: the error pass takes pairs and
    a tail) and an even one: every element within 4 FP32 roundings of the step's increment."""
    res = {}
    for k64 in ["0", "1"]:
        monkeypatch.setenv("MAFX_RK_K64", k64)
        res[k64] = _film_steps(mesh, "f32", "DP45", 1)
    m0, m1_32 = res["0"][0], res["0"][1]
    m1_64 = res["1"][1]
    inc = np.max(np.abs(m1_64 - m0))
    d1 = np.max(np.abs(m1_32 - m1_64))
    print(f"\n[k32 one step] mesh={mesh} increment {inc:.3e} k32 - k64 {d1:.3e} ({d1 / (inc * 2.0 ** -24):.2f} ulps)",
          flush=True)
    assert 0 < d1 <= 4 * 2.0 ** -24 * inc
    assert np.all(np.isfinite(m1_32))


def test_relax_matches_reference_loop():
    """LLGIntegrator::relax (llg_integrator.cpp:43-66) on the GPU against the same loop driven
    through the oracle: E_prev = 1e20; while |(E_prev - E) / E_prev| > precision: E_prev = E,
    ncalcE steps.  Same number of steps, same final state."""
    mesh = (32, 8, 1, 4e-9, 4e-9, 3e-9)
    rng = np.random.default_rng(5)
    m = np.zeros(mesh[:3] + (3,))
    m[..., 0] = 1
    m += 0.3 * rng.standard_normal(m.shape)
    specs = [("demag", {}), ("exchange", {"A": 1.3e-11}), ("zeeman_h", {"h": (0.0, 2e4, 0.0)})]
    state, terms, sim = build(mesh, 8e5, m, specs)
    M = mf()
    llg = M.LLGIntegrator(1.0, terms, "RKF45")
    n_gpu = llg.relax(state, precision=1e-6, ncalcE=10, nprint=0, verbose=False)
    sim.llg_init(1.0, "RKF45")
    E_prev, E = 1e20, sim.energy()[0]
    n_ref = 0
    while abs((E_prev - E) / E_prev) > 1e-6:
        E_prev = E
        for _ in range(10):
            sim.step()
            n_ref += 1
        E = sim.energy()[0]
    assert n_gpu == n_ref and n_gpu > 10
    assert abs(state.t - sim.t) <= 1e-6 * sim.t
    assert np.max(np.abs(state.m - sim.m())) <= 1e-8


@pytest.mark.parametrize("scheme", ["DP45", "RKF45"])
def test_resume_is_step_exact(scheme, tmp_path):
    """save_state after 12 steps, then a fresh State and LLGIntegrator load it and take 12 more:
    the same m to the last bit, the same t and step size as 24 uninterrupted steps."""
    mesh = (32, 8, 1, 4e-9, 4e-9, 3e-9)
    rng = np.random.default_rng(11)
    m = np.zeros(mesh[:3] + (3,))
    m[..., 0] = 1
    m += 0.2 * rng.standard_normal(m.shape)
    M = mf()
    g = M.Mesh(*mesh)

    def terms():
        return [M.DemagField(g), M.ExchangeField(1.3e-11), M.ExternalField(homogeneous=(-2e4, 3e3, 0.0))]

    a = M.State(g, 8e5, m)
    la = M.LLGIntegrator(0.02, terms(), scheme)
    for _ in range(24):
        la.step(a)
    b = M.State(g, 8e5, m)
    lb = M.LLGIntegrator(0.02, terms(), scheme)
    for _ in range(12):
        lb.step(b)
    path = tmp_path / "resume.npz"
    lb.save_state(path, b)
    c = M.State(g, 8e5, np.ones(mesh[:3] + (3,)))
    lc = M.LLGIntegrator(0.02, terms(), scheme)
    lc.load_state(path, c)
    assert c.t == b.t and c.steps == 12
    for _ in range(12):
        lc.step(c)
    assert np.array_equal(c.m.view(np.uint64), a.m.view(np.uint64))
    assert c.t == a.t and lc.controller_state() == la.controller_state()
    # the .vti form of the same state reads back to the last bit
    fn = c.write_vti(tmp_path / "c")
    d = M.State(g, 8e5, np.ones(mesh[:3] + (3,)))
    d.read_vti(fn)
    assert np.array_equal(d.m, c.m)
    # the C++ stream line (state.cpp:26-30) and the Python print (magnumaf.pyx:754-757)
    assert c.stream_line().count("\t") == 3
    assert str(c) == " ".join([str(c.t)] + [str(v) for v in c.mean_m()])


def test_sp4_relaxation_and_switch_gpu():
    # python/test/sp4.py:21-62 on the GPU FP64 path (RKF45, the LLGIntegrator default)
    M = mf()
    nx, ny, nz = 100, 25, 1
    g = M.Mesh(nx, ny, nz, 5.0e-7 / 100, 1.25e-7 / 25, 3.0e-9)
    m = np.zeros((nx, ny, nz, 3))
    m[1:-1, :, :, 0] = 1
    m[0, :, :, 1] = 1
    m[-1, :, :, 1] = 1
    state = M.State(g, 8e5, m)
    llg = M.LLGIntegrator(alpha=1, terms=[M.DemagField(g), M.ExchangeField(1.3e-11)])
    ix = iy = iz = 0.0
    while state.t < 1e-9:
        t1 = state.t
        llg.step(state)
        dt = state.t - t1
        mx, my, mz = state.mean_m()
        ix += mx * dt
        iy += my * dt
        iz += mz * dt
    print(f"\n[sp4 relax] dints = {ix - 9.81206172824e-10:.3e} {iy - 9.14350283169e-11:.3e} "
          f"{iz + 5.74381785359e-13:.3e}", flush=True)
    g = known("sp4_relax")
    assert abs(ix - g["int_mx_dt"]) < g["abs_tol"]
    assert abs(iy - g["int_my_dt"]) < g["abs_tol"]
    assert abs(iz - g["int_mz_dt"]) < g["abs_tol"]
    llg.alpha = 0.02
    z = np.zeros((nx, ny, nz, 3))
    z[..., 0] = -24.6e-3 / MU0
    z[..., 1] = +4.3e-3 / MU0
    llg.add_terms(M.ExternalField(z))
    ix = iy = iz = 0.0
    while state.t < 2e-9:
        t1 = state.t
        llg.step(state)
        dt = state.t - t1
        mx, my, mz = state.mean_m()
        ix += mx * dt
        iy += my * dt
        iz += mz * dt
    print(f"[sp4 switch] dints = {ix + 6.41261165705e-10:.3e} {iy - 1.47353233738e-10:.3e} "
          f"{iz + 1.78144535231e-11:.3e}", flush=True)
    g = known("sp4_switch")
    assert abs(ix - g["int_mx_dt"]) < g["abs_tol"]
    assert abs(iy - g["int_my_dt"]) < g["abs_tol"]
    assert abs(iz - g["int_mz_dt"]) < g["abs_tol"]


def test_integrate_dense_matches_oracle():
    mesh = (32, 8, 1, 4e-9, 4e-9, 3e-9)
    nx, ny, nz = mesh[:3]
    rng = np.random.default_rng(8)
    m = np.zeros((nx, ny, nz, 3))
    m[..., 0] = 1
    m += 0.1 * rng.standard_normal(m.shape)
    h = 25e-3 / MU0
    H = (h * math.cos(math.radians(170)), h * math.sin(math.radians(170)), 0.0)
    specs = [("demag", {}), ("exchange", {"A": 1.3e-11}), ("zeeman_h", {"h": H})]
    state, terms, sim = build(mesh, 8e5, m, specs)
    M = mf()
    llg = M.LLGIntegrator(0.02, terms, "DP45")
    sim.llg_init(0.02, "DP45")
    rows = llg.integrate_dense(state, 5e-11, 1e-11)
    ref = sim.integrate_dense(5e-11, 1e-11)
    assert rows.shape == ref.shape
    assert np.max(np.abs(rows - ref)) <= 1e-10


def test_lbfgs_rows_matches_oracle():
    """LBFGS on rows of 256 cells, where the gradient, its maxima and the energies come from the
    z-column lanes with per-group partials: the final energy of 60 iterations against the oracle's
    (lbfgs_minimizer.cpp:55-198)."""
    mesh = (256, 4, 2, 3e-9, 3e-9, 3e-9)
    m = rand_m(mesh[:3], 23)
    specs = [("demag", {}), ("exchange", {"A": 1.5e-11}), ("dmi", {"D": 3e-3}), ("uniaxial", {"K": 8e5}),
             ("zeeman_h", {"h": (0, 0, 0.2 / MU0)})]
    state, terms, sim = build(mesh, 1.1e6, m, specs)
    M = mf()
    E0 = M.LLGIntegrator(0.1, terms).Eeff_in_J(state)
    mini = M.LBFGS_Minimizer(terms, tol=1e-6, maxiter=60)
    E = mini.minimize(state)
    r = sim.lbfgs(tol=1e-6, maxiter=60)
    print(f"\n[lbfgs rows] E0={E0:.6e} gpu E={E:.12e} it={mini.last_stats['iterations']} "
          f"oracle E={r['E']:.12e} it={r['iterations']}", flush=True)
    assert E < E0
    assert abs(E - r["E"]) <= 1e-9 * abs(r["E"])


def test_lbfgs_matches_oracle():
    mesh = (16, 16, 1, 3e-9, 3e-9, 3e-9)
    m = rand_m(mesh[:3], 21)
    specs = [("demag", {}), ("exchange", {"A": 1.5e-11}), ("dmi", {"D": 3e-3}), ("uniaxial", {"K": 8e5}),
             ("zeeman_h", {"h": (0, 0, 0.2 / MU0)})]
    state, terms, sim = build(mesh, 1.1e6, m, specs)
    M = mf()
    E0 = M.LLGIntegrator(0.1, terms).Eeff_in_J(state)
    mini = M.LBFGS_Minimizer(terms, tol=1e-6, maxiter=60)
    E = mini.minimize(state)
    r = sim.lbfgs(tol=1e-6, maxiter=60)
    print(f"\n[lbfgs] E0={E0:.6e} gpu E={E:.12e} it={mini.last_stats['iterations']} "
          f"oracle E={r['E']:.12e} it={r['iterations']}", flush=True)
    assert E < E0
    assert abs(E - r["E"]) <= 1e-9 * abs(r["E"])
