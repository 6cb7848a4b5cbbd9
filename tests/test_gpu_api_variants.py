"""GPU parity of the reference's spatially varying and time-dependent API variants against the FP64
oracle, and of parameter changes between evaluations:

- DmiField(af::array D_constants, axis)                 dmi_field.cpp:17-19,97-109
- UniaxialAnisotropyField(Ku1_field | Ku1, axis | Ku1_axis_field)
                                                        uniaxial_anisotropy_field.cpp:11-25,58-89
- LLGIntegrator(af::array alpha_field, ...)            llg_integrator.hpp:32-40, llg_integrator.cpp:18-38
- ExternalField(callback) evaluated at the stage times  external_field.cpp:13-17,28-36,
                                                        adaptive_runge_kutta.cpp:113-173
- ExternalField::set_homogeneous_field, term members and State.Ms changed between evaluations
                                                        external_field.cpp:21-26, magnumaf.pyx:777-784

Tolerances as in test_gpu_parity.py: FP64 H_eff normwise <= 1e-12, energies <= 1e-12 of the energy
scale, integrator runs with identical accept/reject decisions, h within 1e-6 and m within 1e-8."""
import math

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

MU0 = O.MU0
FP64_TOL = 1e-12
FP32_TOL = 1e-5


def mf():
    import magnumaf
    return magnumaf


def normwise(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def rand_unit(shape, seed):
    m = np.random.default_rng(seed).standard_normal(tuple(shape) + (3,))
    return m / np.linalg.norm(m, axis=-1, keepdims=True)


def check_field_and_energy(state, gterm, sim, precision="f64"):
    h = gterm.H_in_Apm(state, precision=precision)
    ref = sim.heff()
    err = normwise(h, ref)
    E = gterm.Energy_in_J(state, precision=precision)
    Eref = sim.energy()[0]
    return err, abs(E - Eref) / max(abs(Eref), 1e-300)


@pytest.mark.parametrize("mesh", [(16, 8, 4, 2e-9, 2e-9, 2e-9), (33, 7, 1, 1e-9, 2e-9, 1.5e-9)])
@pytest.mark.parametrize("ms_field", [False, True])
def test_dmi_D_field(mesh, ms_field):
    print(f"\n[dmi D field] mesh={mesh} ms_field={ms_field}", flush=True)
    M = mf()
    nx, ny, nz = mesh[:3]
    rng = np.random.default_rng(3)
    m = rand_unit((nx, ny, nz), 4)
    D = rng.uniform(1e-3, 4e-3, (nx, ny, nz))
    D[0, 0, 0] = 0.0
    Ms = 1.1e6
    Msf = None
    if ms_field:
        Msf = rng.uniform(5e5, 1.2e6, (nx, ny, nz))
        Msf[1, 0, 0] = 0.0
    g = M.Mesh(*mesh)
    state = M.State(g, Ms if Msf is None else Msf, m)
    term = M.DmiField(D, [0.2, 0.1, 1.0])
    sim = O.Sim(mesh, Ms, m, [O.term(O.DMI, 0.0, axis=(0.2, 0.1, 1.0), array0=O.scalar_flat(D))],
                Ms_field=None if Msf is None else Msf)
    err, eerr = check_field_and_energy(state, term, sim)
    assert err <= FP64_TOL and eerr <= FP64_TOL, (err, eerr)
    err32, _ = check_field_and_energy(state, term, sim, precision="f32")
    assert err32 <= FP32_TOL, err32


@pytest.mark.parametrize("variant", ["Ku1_field", "axis_field", "both"])
@pytest.mark.parametrize("ms_field", [False, True])
def test_uniaxial_fields(variant, ms_field):
    mesh = (12, 10, 3, 2e-9, 2e-9, 2e-9)
    print(f"\n[uniaxial] {variant} ms_field={ms_field}", flush=True)
    M = mf()
    nx, ny, nz = mesh[:3]
    rng = np.random.default_rng(7)
    m = rand_unit((nx, ny, nz), 8)
    K = rng.uniform(2e5, 9e5, (nx, ny, nz))
    ax = rng.standard_normal((nx, ny, nz, 3))  # normalised per cell by the term
    Ms = 8e5
    Msf = None
    if ms_field:
        Msf = rng.uniform(5e5, 1.2e6, (nx, ny, nz))
        Msf[2, 3, 1] = 0.0
    Kg = K if variant in ("Ku1_field", "both") else 4e5
    axg = ax if variant in ("axis_field", "both") else (0.3, 0.0, 1.0)
    g = M.Mesh(*mesh)
    state = M.State(g, Ms if Msf is None else Msf, m)
    term = M.UniaxialAnisotropyField(Kg, axg)
    sim = O.Sim(mesh, Ms, m, [O.term(O.UNIAXIAL, 0.0 if variant != "axis_field" else 4e5,
                                     axis=(0.3, 0.0, 1.0) if variant == "Ku1_field" else (0.0, 0.0, 1.0),
                                     array0=O.scalar_flat(K) if variant in ("Ku1_field", "both") else None,
                                     array1=O.to_flat(ax) if variant in ("axis_field", "both") else None)],
                Ms_field=None if Msf is None else Msf)
    err, eerr = check_field_and_energy(state, term, sim)
    assert err <= FP64_TOL and eerr <= FP64_TOL, (err, eerr)
    # the reference's 2 / mu0 per cell with an axis field along m (python/test/micro_anisotropy.py:36-54)
    if variant == "both" and not ms_field:
        u = ax / np.linalg.norm(ax, axis=-1, keepdims=True)
        st2 = M.State(g, Ms, u)
        h = term.H_in_Apm(st2)
        assert np.allclose(h, 2 * K[..., None] / (MU0 * Ms) * u, rtol=1e-14, atol=0)


def test_llg_alpha_field_replay():
    """LLGIntegrator(alpha_field) (llg_integrator.cpp:18-38): DP45 replay against the oracle, then
    the dissipation-only right-hand side with the same field."""
    mesh = (32, 8, 1, 4e-9, 4e-9, 3e-9)
    print(f"\n[alpha field] mesh={mesh}", flush=True)
    M = mf()
    nx, ny, nz = mesh[:3]
    rng = np.random.default_rng(11)
    m = np.zeros((nx, ny, nz, 3))
    m[..., 0] = 1
    m += 0.2 * rng.standard_normal(m.shape)
    alpha = rng.uniform(0.01, 0.5, (nx, ny, nz))
    H = (2e4, -1e4, 5e3)
    g = M.Mesh(*mesh)
    for diss in (False, True):
        state = M.State(g, 8e5, m)
        terms = [M.DemagField(g), M.ExchangeField(1.3e-11), M.ExternalField(homogeneous=H)]
        llg = M.LLGIntegrator(alpha, terms, "DP45", dissipation_term_only=diss)
        sim = O.Sim(mesh, 8e5, m, [O.term(O.DEMAG), O.term(O.EXCHANGE, 1.3e-11), O.term(O.ZEEMAN, field=H)])
        sim.llg_init(0.0, "DP45", dissipation_term_only=diss)
        sim.set_alpha_field(alpha)
        for i in range(25):
            r = sim.step()
            llg.step(state)
            st = llg.stats()
            assert abs(st["h_last"] - r["h"]) <= 1e-6 * r["h"], (diss, i, st["h_last"], r["h"])
            assert st["rejected"] == r["rejected"] and st["accepted"] == r["accepted"]
        assert abs(state.t - sim.t) <= 1e-6 * sim.t
        assert np.max(np.abs(state.m - sim.m())) <= 1e-8
        assert np.array_equal(llg.alpha, alpha)
    # back to a scalar alpha between steps (the alpha setter of magnumaf.pyx:1010-1014)
    state = M.State(g, 8e5, m)
    llg = M.LLGIntegrator(alpha, [M.DemagField(g), M.ExchangeField(1.3e-11)], "DP45")
    sim = O.Sim(mesh, 8e5, m, [O.term(O.DEMAG), O.term(O.EXCHANGE, 1.3e-11)])
    sim.llg_init(0.0, "DP45")
    sim.set_alpha_field(alpha)
    for i in range(6):
        if i == 3:
            llg.alpha = 0.1
            sim.set_alpha_field(None)
            sim.set_alpha(0.1)
        r = sim.step()
        llg.step(state)
        assert abs(llg.stats()["h_last"] - r["h"]) <= 1e-6 * r["h"], i
    assert llg.alpha == 0.1
    assert np.max(np.abs(state.m - sim.m())) <= 1e-8


def _rotating(t):
    w = 2 * math.pi * 5e9
    return (3e4 * math.cos(w * t), 3e4 * math.sin(w * t), 1e4)


@pytest.mark.parametrize("mode", ["homogeneous", "full"])
def test_time_dependent_external_field(mode):
    """ExternalField(callback) (external_field.cpp:28-36): the field is evaluated at the State's
    time, and inside an RK attempt at the stage times t + c_i h (adaptive_runge_kutta.cpp:113-173).
    A 5 GHz rotating field; DP45 replay against the oracle with the same callback."""
    mesh = (16, 8, 2, 4e-9, 4e-9, 3e-9)
    print(f"\n[time-dependent field] {mode}", flush=True)
    M = mf()
    nx, ny, nz = mesh[:3]
    m = rand_unit((nx, ny, nz), 12)
    xs = (np.arange(nx) + 0.5) / nx

    def full_np(t):  # (nx, ny, nz, 3): the rotating field with an x ramp
        h = np.array(_rotating(t))
        f = np.empty((nx, ny, nz, 3))
        f[...] = h
        f[..., 0] *= xs[:, None, None]
        return f

    g = M.Mesh(*mesh)
    if mode == "homogeneous":
        gz = M.ExternalField(lambda st: _rotating(st.t), homogeneous_callback=True)
        oz = O.term(O.ZEEMAN, callback=lambda t, mm: _rotating(t), homogeneous_callback=True)
    else:
        gz = M.ExternalField(lambda st: full_np(st.t))
        oz = O.term(O.ZEEMAN, callback=lambda t, mm: O.to_flat(full_np(t)))
    state = M.State(g, 8e5, m)
    terms = [M.ExchangeField(1.3e-11), gz]
    sim = O.Sim(mesh, 8e5, m, [O.term(O.EXCHANGE, 1.3e-11), oz])
    # the field at a set time (State.t), and its Zeeman energy
    state.t = 37e-12
    sim.t = 37e-12
    h = gz.H_in_Apm(state)
    expect = full_np(37e-12) if mode == "full" else np.broadcast_to(np.array(_rotating(37e-12)), h.shape)
    assert normwise(h, expect) <= 1e-15
    E, Eref = gz.Energy_in_J(state), sim.energy(mask=2)[0]
    assert abs(E - Eref) <= 1e-12 * abs(Eref)
    llg = M.LLGIntegrator(0.05, terms, "DP45")
    sim.llg_init(0.05, "DP45")
    for i in range(30):
        r = sim.step()
        llg.step(state)
        st = llg.stats()
        assert abs(st["h_last"] - r["h"]) <= 1e-6 * r["h"], (i, st["h_last"], r["h"])
        assert st["rejected"] == r["rejected"] and st["accepted"] == r["accepted"]
    assert abs(state.t - sim.t) <= 1e-6 * sim.t
    assert np.max(np.abs(state.m - sim.m())) <= 1e-8


def test_live_parameter_changes():
    """Changes between evaluations reach the next one, as in the reference's hysteresis loops
    (python/examples/loop_external_field.py): set_homogeneous_field in a loop, a term member, and
    State.Ms, each checked against an oracle built with the new parameters, for H, the energy and
    one LLG step."""
    mesh = (16, 8, 2, 3e-9, 3e-9, 3e-9)
    print("\n[live parameter changes]", flush=True)
    M = mf()
    nx, ny, nz = mesh[:3]
    m = rand_unit((nx, ny, nz), 21)
    g = M.Mesh(*mesh)
    state = M.State(g, 8e5, m)
    zee = M.ExternalField(homogeneous=(0.0, 0.0, 0.0))
    uni = M.UniaxialAnisotropyField(2e5, (0, 0, 1))
    terms = [M.DemagField(g), M.ExchangeField(1.3e-11), uni, zee]
    llg = M.LLGIntegrator(0.1, terms, "RKF45")

    def oracle(Ms, K, H, mm):
        return O.Sim(mesh, Ms, mm, [O.term(O.DEMAG), O.term(O.EXCHANGE, 1.3e-11), O.term(O.UNIAXIAL, K),
                                    O.term(O.ZEEMAN, field=H)])

    for k, hz in enumerate([-5e4, 0.0, 2e4, 8e4]):
        zee.set_homogeneous_field(1e3 * k, 0.0, hz)
        mm = state.m
        sim = oracle(8e5, 2e5, (1e3 * k, 0.0, hz), mm)
        assert normwise(llg.H_in_Apm(state), sim.heff()) <= FP64_TOL
        E, Eref = llg.Eeff_in_J(state), sim.energy()[0]
        assert abs(E - Eref) <= FP64_TOL * max(abs(Eref), MU0 * 8e5 ** 2 * g.n * 27e-27)
    uni.Ku1 = 6e5
    state.Ms = 1.0e6
    mm = state.m
    sim = oracle(1.0e6, 6e5, (3e3, 0.0, 8e4), mm)
    assert normwise(llg.H_in_Apm(state), sim.heff()) <= FP64_TOL
    # steps of the integrator built before the changes, against the oracle with the new parameters
    sim.llg_init(0.1, "RKF45")
    for i in range(5):
        r = sim.step()
        llg.step(state)
        assert abs(llg.stats()["h_last"] - r["h"]) <= 1e-6 * r["h"], i
    assert np.max(np.abs(state.m - sim.m())) <= 1e-8


@pytest.mark.parametrize("precision", ["f32", "f32_demag", "f32_stencils"])
def test_mixed_precision_modes(precision):
    """MAFX_F32_DEMAG (FP32 demag, FP64 local terms) and MAFX_F32_STENCILS (the reverse) against the
    oracle, with the FP32-field bound; each term alone shows which part runs in FP32: the FP64 part
    matches the oracle to the FP64 bound."""
    mesh = (64, 32, 4, 2e-9, 2e-9, 2e-9)
    print(f"\n[precision] {precision}", flush=True)
    M = mf()
    m = rand_unit(mesh[:3], 31)
    g = M.Mesh(*mesh)
    state = M.State(g, 1.1e6, m)
    terms = [M.DemagField(g), M.ExchangeField(1.5e-11), M.DmiField(3e-3), M.UniaxialAnisotropyField(8e5)]
    sim = O.Sim(mesh, 1.1e6, m, [O.term(O.DEMAG), O.term(O.EXCHANGE, 1.5e-11), O.term(O.DMI, 3e-3),
                                 O.term(O.UNIAXIAL, 8e5)])
    ctx = state._context(terms, precision)
    errs = []
    for k in range(4):
        h = ctx.heff(state.m_dev, mask=1 << k).cpu().numpy()
        errs.append(normwise(h, sim.heff_flat(1 << k)))
    h = ctx.heff(state.m_dev).cpu().numpy()
    assert normwise(h, sim.heff_flat()) <= FP32_TOL
    # MAFX_F32_STENCILS: the FP64 demag field is only rounded to FP32 where the per-cell kernel adds
    # it to the FP32 local terms (2^-24 relative); MAFX_F32_DEMAG: the local terms are FP64
    if precision == "f32_stencils":
        assert errs[0] <= 2e-7, errs
    else:
        assert FP64_TOL < errs[0] <= FP32_TOL, errs
    if precision == "f32_demag":
        assert all(e <= FP64_TOL for e in errs[1:]), errs
    else:
        assert all(FP64_TOL < e <= FP32_TOL for e in errs[1:]), errs


def test_fp32_demag_far_field_512x512x2():
    """The FP32 demag field where the far field carries weight (512 x 512 x 2 cells of 2 nm, padded
    1024 x 1024 x 4), against the oracle.  The Newell tensor's far-field entries are cancellation
    noise of the 27-term second difference (demag_field.cpp:71-102): any other evaluation of the same
    formulas (the FP32 path's parity-extended octant, the GPU's libm) gives a field that differs from
    the reference's by that noise, while the FP32 arithmetic alone, measured against the FP64 pipeline
    on the FP32 path's own tensor, stays at the 1e-6 level.  Bounds from the measured values at this
    size (printed), with headroom."""
    mesh = (512, 512, 2, 2e-9, 2e-9, 2e-9)
    print("\n[fp32 demag far field] 512x512x2", flush=True)
    M = mf()
    m = rand_unit(mesh[:3], 41)
    g = M.Mesh(*mesh)
    state = M.State(g, 1.1e6, m)
    ref = O.Sim(mesh, 1.1e6, m, [O.term(O.DEMAG)]).heff_flat()

    def field(prec, **kw):
        return state._context([M.DemagField(g, **kw)], prec).heff(state.m_dev).cpu().numpy()

    h64 = field("f64")
    h64_oct = field("f64", octant_tensor=True)
    h32 = field("f32")
    h64_dev = field("f64", device_tensor=True)
    h32_full = field("f32", full_tensor=True)
    e = {"fp64": normwise(h64, ref), "fp64_octant": normwise(h64_oct, ref), "fp64_gpu_tensor": normwise(h64_dev, ref),
         "fp32_vs_fp64_octant": normwise(h32, h64_oct), "fp32": normwise(h32, ref),
         "fp32_full_tensor": normwise(h32_full, ref)}
    print(f"[fp32 demag far field] {e}", flush=True)
    assert e["fp64"] <= FP64_TOL
    assert e["fp32_vs_fp64_octant"] <= 1e-5          # FP32 arithmetic alone
    assert e["fp32"] <= 1e-3 and e["fp64_octant"] <= 1e-3  # tensor noise of another evaluation order
    # on the reference's own tensor spectrum (MAFX_TERM_FULL_TENSOR) only the FP32 arithmetic is left
    assert e["fp32_full_tensor"] <= 1e-5
