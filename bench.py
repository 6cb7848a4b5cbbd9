"""Benchmark of the effective-field / integrator hot path on one MI355X.

Workload (default): the 1024 x 1024 x 4 film of 2 nm cubes, Ms = 1.1e6 A/m, A = 1.5e-11 J/m,
uniaxial Ku = 8e5 J/m^3 along z, interface DMI D = 3e-3 J/m^2, demag, alpha = 0.1, no external
field, seeded uniform random unit vectors as initial m.

A "step" of the timed region is one H_eff evaluation of the FP32-field path (demag FFT + tensor
product + stencils in FP32, m in FP64) with all inputs resident in HBM; `value` is H_eff
evaluations per second.  The JSON line also carries the FP64-path H_eff rate, DP45 steps/s of both
paths, per-kernel HIP-event timings, the roofline of the dominant kernel, and the restated
reference (the oracle, FP64 on the host cores) as the CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--dp45-steps S]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "magnum.af_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
MU0 = 4e-7 * math.pi


def workload(nx, ny, nz, seed=2024):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((nx, ny, nz, 3))
    m /= np.linalg.norm(m, axis=-1, keepdims=True)
    return m


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dp45-steps", type=int, default=12)
    ap.add_argument("--size", default="1024x1024x4")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--heff-only", choices=["f32", "f64"], default=None,
                    help="run --steps H_eff evaluations of one path and exit (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (the sharding design is "
              "described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import magnumaf as M

    nx, ny, nz = (int(v) for v in args.size.split("x"))
    d = 2e-9
    Ms, A, Ku, D, alpha = 1.1e6, 1.5e-11, 8e5, 3e-3, 0.1
    mesh = M.Mesh(nx, ny, nz, d, d, d)
    m0 = workload(nx, ny, nz)
    terms = [M.DemagField(mesh), M.ExchangeField(A), M.DmiField(D, [0, 0, 1]), M.UniaxialAnisotropyField(Ku, [0, 0, 1])]

    t0 = time.time()
    state = M.State(mesh, Ms, m0)
    if args.heff_only:
        ctx = state._context(terms, args.heff_only)
        h = torch.empty(3 * mesh.n, dtype=torch.float64, device="cuda")
        for _ in range(args.warmup + args.steps):
            ctx.heff(state.m_dev, out=h)
        torch.cuda.synchronize()
        log(f"[bench] {args.steps} {args.heff_only} H_eff evaluations done ({time.time() - t0:.1f} s)")
        return
    ctx32 = state._context(terms, "f32")
    log(f"[bench] FP32 context ready in {time.time() - t0:.1f} s (includes host Newell tensor)")
    t0 = time.time()
    ctx64 = state._context(terms, "f64")
    log(f"[bench] FP64 context ready in {time.time() - t0:.1f} s")
    N = mesh.n
    k32 = torch.empty(3 * N, dtype=torch.float64, device="cuda")

    def heff_loop(ctx, n):
        for _ in range(n):
            ctx.heff(state.m_dev, out=k32)

    # ---------------------------------------------------------------- timed region: FP32 H_eff
    heff_loop(ctx32, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    heff_loop(ctx32, args.steps)
    torch.cuda.synchronize()
    t_fp32 = time.perf_counter() - t0
    heff32 = args.steps / t_fp32
    log(f"[bench] FP32-field H_eff: {heff32:.1f} /s ({1e3 * t_fp32 / args.steps:.3f} ms)")

    # ---------------------------------------------------------------- FP64 H_eff
    k64n = max(20, args.steps // 4)
    heff_loop(ctx64, 3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    heff_loop(ctx64, k64n)
    torch.cuda.synchronize()
    heff64 = k64n / (time.perf_counter() - t0)
    log(f"[bench] FP64 H_eff: {heff64:.1f} /s")

    # ---------------------------------------------------------------- per-kernel HIP-event timing
    def kernel_table(ctx, n):
        ctx.prof_reset()
        ctx.prof(True)
        heff_loop(ctx, n)
        rows = ctx.prof_report()
        ctx.prof(False)
        out = {}
        for name, r in rows.items():
            avg_ms = r["ms"] / r["count"]
            out[name] = dict(launches=r["count"], avg_ms=avg_ms, bytes=r["bytes"],
                             GBps=r["bytes"] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0)
        return out

    kt32 = kernel_table(ctx32, 20)
    kt64 = kernel_table(ctx64, 10)
    dom = max(kt32.items(), key=lambda kv: kv[1]["avg_ms"] * kv[1]["launches"])
    per_heff_ms32 = sum(v["avg_ms"] * v["launches"] for v in kt32.values()) / 20
    bytes_heff32 = sum(v["bytes"] * v["launches"] for v in kt32.values()) / 20
    per_heff_ms64 = sum(v["avg_ms"] * v["launches"] for v in kt64.values()) / 10
    bytes_heff64 = sum(v["bytes"] * v["launches"] for v in kt64.values()) / 10

    # plain copy bandwidth as the achievable ceiling (read + write of 1 GiB)
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    copy_gbs = 10 * 2 * src.numel() * 8 / (ev0.elapsed_time(ev1) * 1e-3) / 1e9
    del src, dst

    # ---------------------------------------------------------------- DP45 steps/s
    def dp45_rate(precision, nsteps):
        st = M.State(mesh, Ms, m0)
        llg = M.LLGIntegrator(alpha, terms, "DP45", precision=precision)
        st._ctx = state._ctx  # reuse the contexts built above
        for _ in range(3):
            llg.step(st)
        torch.cuda.synchronize()
        s0 = llg.stats()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            llg.step(st)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        s1 = llg.stats()
        attempts = (s1["heff_evals"] - s0["heff_evals"]) / 7
        return nsteps / dt, attempts / dt, s1

    dp32, at32, s32 = dp45_rate("f32", args.dp45_steps)
    log(f"[bench] DP45 FP32-field: {dp32:.2f} steps/s ({at32:.2f} attempts/s)")
    dp64, at64, s64 = dp45_rate("f64", max(4, args.dp45_steps // 2))
    log(f"[bench] DP45 FP64: {dp64:.2f} steps/s ({at64:.2f} attempts/s)")

    # ---------------------------------------------------------------- CPU baseline (oracle)
    cpu = None
    if not args.no_cpu:
        import oracle_lib as O
        threads = int(os.environ.get("OMP_NUM_THREADS", str(os.cpu_count() or 1)))
        t0 = time.time()
        sim = O.Sim((nx, ny, nz, d, d, d), Ms, m0,
                    [O.term(O.DEMAG), O.term(O.EXCHANGE, A), O.term(O.DMI, D), O.term(O.UNIAXIAL, Ku)],
                    nthreads=threads)
        log(f"[bench] oracle setup {time.time() - t0:.1f} s")
        t0 = time.perf_counter()
        sim.heff_flat()
        cpu_s = time.perf_counter() - t0
        cpu = {"value": 1.0 / cpu_s, "unit": "H_eff evals/s", "cores": threads, "kind": "port",
               "sample": f"one FP64 H_eff of the restated reference (zero-padded FFT demag + 27-tap stencils) "
                         f"at the full {nx}x{ny}x{nz} mesh on {threads} host threads; tensor setup untimed"}
        log(f"[bench] oracle H_eff {cpu_s:.2f} s")

    out = {
        "metric": "H_eff evals/s & DP45 steps/s at 1024^2x4 cells; LBFGS time-to-tolerance",
        "value": heff32,
        "unit": "H_eff evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_fp32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 field / f64 state",
        "data": "synthetic (seeded random unit vectors)",
        "config": {"workload": f"C2 film {nx}x{ny}x{nz} cells of 2 nm, demag+exchange+DMI+uniaxial, alpha=0.1",
                   "cells": N, "path": "FP32-field H_eff"},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": dom[1]["GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom[1]["GBps"] / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
        "extra": {
            "heff_per_s_fp32": heff32, "heff_per_s_fp64": heff64,
            "dp45_steps_per_s_fp32": dp32, "dp45_steps_per_s_fp64": dp64,
            "dp45_attempts_per_s_fp32": at32, "dp45_attempts_per_s_fp64": at64,
            "dp45_fp32_over_fp64": dp32 / dp64 if dp64 > 0 else None,
            "heff_algorithmic_GB_fp32": bytes_heff32 / 1e9, "heff_algorithmic_GB_fp64": bytes_heff64 / 1e9,
            "heff_kernel_ms_fp32": per_heff_ms32, "heff_kernel_ms_fp64": per_heff_ms64,
            "heff_frac_of_hbm_peak_fp32": bytes_heff32 / (per_heff_ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "heff_frac_of_hbm_peak_fp64": bytes_heff64 / (per_heff_ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "copy_kernel_GBps": copy_gbs,
            "kernels_fp32": kt32, "kernels_fp64": kt64,
        },
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
