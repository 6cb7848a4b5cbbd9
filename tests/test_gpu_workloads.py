"""GPU runs of the C3, C4 and C5 workloads against the oracle at sizes the oracle finishes in seconds.

- C3: LBFGS (lbfgs_minimizer.cpp:55-505) on the atomistic term set (src/field_terms/atom/*.cpp) of a
  32 x 32 x 1 spin film: the final energy within 1e-9 and the iteration count within 5% (the
  dipole FFT's rounding differs between the two, and the triple stop test then fires a few
  iterations apart on the flat tail of the minimisation).
- C5: the multilayer term list (demag, RKKY with rkky = J dz / 2 between layers as in
  test/rkky.cpp:19,28, interface DMI, uniaxial anisotropy, in-plane field) on a 256 x 4 x 8 mesh:
  DP45 step replay on the FP64 path; the FP32-field path's H_eff within its 1e-5 bound; FP32-field
  LBFGS ending at the FP64 LBFGS energy within the FP32 field's accuracy.
- C4: warm start -> LBFGS on a 64 x 64 film (the C4 term set) with a network trained briefly on the
  GPU, judged by plain LBFGS: both reach max|m x H| / max|H| <= 1e-6, and the warm start's final
  energy is either equal (<= 1e-9) or the run is counted as a different minimum.
"""
import math

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

MU0 = O.MU0
MU_B = 9.274009994e-24


def mf():
    import magnumaf
    return magnumaf


def rand_unit(shape, seed):
    m = np.random.default_rng(seed).standard_normal(tuple(shape) + (3,))
    return m / np.linalg.norm(m, axis=-1, keepdims=True)


def test_c3_atomistic_lbfgs():
    """C3's term set (exchange J, anisotropy 0.02 J, interface DMI 0.2 J with axis (0, 0, -1),
    dipole-dipole, Bz = 1 T) on 32 x 32 x 1 spins of a = 0.27 nm from random spins."""
    M = mf()
    mesh = (32, 32, 1, 0.27e-9, 0.27e-9, 0.27e-9)
    J, mus = 1.6e-21, 2.6 * MU_B
    m = rand_unit(mesh[:3], 3)
    g = M.Mesh(*mesh)
    terms = [M.AtomisticExchangeField(J), M.AtomisticUniaxialAnisotropyField(0.02 * J, (0, 0, 1)),
             M.AtomisticDmiField(0.2 * J, (0, 0, -1)), M.AtomisticDipoleDipoleField(g),
             M.AtomisticExternalField(homogeneous=(0, 0, 1.0 / MU0))]
    oterms = [O.term(O.ATOM_EXCHANGE, J), O.term(O.ATOM_UNIAXIAL, 0.02 * J), O.term(O.ATOM_DMI, 0.2 * J, axis=(0, 0, -1)),
              O.term(O.ATOM_DIPOLE), O.term(O.ATOM_ZEEMAN, field=(0, 0, 1.0 / MU0))]
    state = M.State(g, mus, m)
    sim = O.Sim(mesh, mus, m, oterms)
    mini = M.LBFGS_Minimizer(terms, tol=1e-12, maxiter=400)
    E = mini.minimize(state)
    r = sim.lbfgs(tol=1e-12, maxiter=400)
    st = mini.last_stats
    print(f"\n[C3 32x32] gpu E={E:.12e} it={st['iterations']}  oracle E={r['E']:.12e} it={r['iterations']}", flush=True)
    assert abs(st["iterations"] - r["iterations"]) <= 0.05 * r["iterations"]
    assert abs(E - r["E"]) <= 1e-9 * abs(r["E"])


def _c5_terms(M, g, nx, ny, nz, d):
    rk = np.full((nx, ny, nz), -1e-3 * d / 2.0)
    ex = np.full((nx, ny, nz), 1.3e-11)
    H = (50e-3 / MU0, 0.0, 0.0)
    gterms = [M.DemagField(g), M.RKKYExchangeField(rk, ex, g), M.DmiField(2.5e-3, [0, 0, 1]),
              M.UniaxialAnisotropyField(7e5, [0, 0, 1]), M.ExternalField(homogeneous=H)]
    oterms = [O.term(O.DEMAG), O.term(O.RKKY, array0=O.scalar_flat(rk), array1=O.scalar_flat(ex)),
              O.term(O.DMI, 2.5e-3), O.term(O.UNIAXIAL, 7e5), O.term(O.ZEEMAN, field=H)]
    return gterms, oterms


def test_c5_multilayer_dp45_replay_and_fp32():
    M = mf()
    nx, ny, nz, d = 256, 4, 8, 2e-9
    mesh = (nx, ny, nz, d, d, d)
    g = M.Mesh(*mesh)
    m = rand_unit((nx, ny, nz), 5)
    gterms, oterms = _c5_terms(M, g, nx, ny, nz, d)
    state = M.State(g, 1.0e6, m)
    sim = O.Sim(mesh, 1.0e6, m, oterms)
    ref = sim.heff()
    h64 = M.LLGIntegrator(0.05, gterms).H_in_Apm(state)
    h32 = M.LLGIntegrator(0.05, gterms, precision="f32").H_in_Apm(state)
    e64 = np.max(np.abs(h64 - ref)) / np.max(np.abs(ref))
    e32 = np.max(np.abs(h32 - ref)) / np.max(np.abs(ref))
    print(f"\n[C5 256x4x8] H_eff FP64 {e64:.2e}, FP32 {e32:.2e}", flush=True)
    assert e64 <= 1e-12 and e32 <= 1e-5
    llg = M.LLGIntegrator(0.05, gterms, "DP45")
    sim.llg_init(0.05, "DP45")
    for i in range(30):
        r = sim.step()
        llg.step(state)
        st = llg.stats()
        assert abs(st["h_last"] - r["h"]) <= 1e-6 * r["h"], (i, st["h_last"], r["h"])
        assert st["rejected"] == r["rejected"] and st["accepted"] == r["accepted"]
    # from random spins the multilayer amplifies the ~1e-16 differences of the two FFTs over 30
    # steps to ~4e-8 (the FP64 films of test_gpu_parity.py start near +x and stay within 1e-8)
    assert np.max(np.abs(state.m - sim.m())) <= 2e-7


def test_c5_multilayer_fp32_lbfgs():
    """FP32-field LBFGS against FP64 LBFGS on the C5 term list from the same start, both to
    max|m x H| / max|H| <= 1e-6 (bounded iterations): the FP32 field's own accuracy (~1e-6 of the
    largest term) bounds how far its minimum can sit from the FP64 one.  The start, +z with 20%
    noise, lies in one basin; from fully random spins the two rounding paths may end in different
    metastable states of the multilayer."""
    M = mf()
    nx, ny, nz, d = 64, 16, 8, 2e-9
    g = M.Mesh(nx, ny, nz, d, d, d)
    m = 0.2 * np.random.default_rng(6).standard_normal((nx, ny, nz, 3))
    m[..., 2] += 1.0
    gterms, _ = _c5_terms(M, g, nx, ny, nz, d)
    res = {}
    for prec in ["f64", "f32"]:
        state = M.State(g, 1.0e6, m)
        mini = M.LBFGS_Minimizer(gterms, tol=1e-20, maxiter=3000, torque_tol=1e-6, precision=prec)
        E = mini.minimize(state)
        E64 = M.LLGIntegrator(0.05, gterms).Eeff_in_J(state)  # energy of the end state on the FP64 path
        res[prec] = (E, E64, mini.last_stats, state.m)
        print(f"\n[C5 LBFGS {prec}] E={E:.10e} E(fp64 eval)={E64:.10e} {mini.last_stats}", flush=True)
    s64, s32 = res["f64"][2], res["f32"][2]
    # both end on the torque stop; the FP32 run hands over to FP64 for its last iterations when
    # its energy noise stalls it first, and does most of its iterations on the FP32 path
    assert s64["stop_reason"] == 3 and s32["stop_reason"] == 3
    assert s32.get("fp32_iterations", s32["iterations"]) >= s32.get("fp64_iterations", 0)
    assert abs(res["f32"][1] - res["f64"][1]) <= 1e-7 * abs(res["f64"][1])


def test_c4_warm_start_then_lbfgs():
    """C4's pipeline on 64 x 64 films: FP64 LBFGS targets for a handful of training films, a short
    training of the encoder-decoder on the GPU, then warm start + LBFGS against plain LBFGS from the
    same random start.  Both must reach the torque tolerance; the run reports iterations and whether
    the final energies agree within 1e-9."""
    import magnumaf.warmstart as WS
    data = WS.generate_dataset(8, seed=1, size=64, torque_tol=1e-4)
    net = WS.WarmStartNet()
    hist, hold = WS.train(net, data, epochs=3, batch=4, lr=1e-3)
    assert all(np.isfinite(hist))
    ev = WS.evaluate(net, n=2, size=64, seed=77)
    for e in ev:
        print(f"\n[C4 64x64] {e}", flush=True)
        assert e["stop_plain"] == 3 and e["stop_warm"] == 3  # both stopped on the torque tolerance
        assert e["it_plain"] > 0 and e["it_warm"] > 0
