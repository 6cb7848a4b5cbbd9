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
COPY_CAL_BYTES = 1 << 28  # calibration copy of the profiler runs: 256 MiB read + 256 MiB written


def workload(nx, ny, nz, seed=2024):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((nx, ny, nz, 3))
    m /= np.linalg.norm(m, axis=-1, keepdims=True)
    return m


def log(msg):
    print(msg, file=sys.stderr, flush=True)


MIN_TIMED_S = 0.5


def timed_repeats(run, torch, min_s=MIN_TIMED_S, min_reps=3, max_reps=1000):
    """Wall seconds of repeated calls of run(), each bracketed by device synchronisations, until
    min_s seconds in total (and min_reps repeats) have been timed."""
    reps = []
    while len(reps) < min_reps or (sum(reps) < min_s and len(reps) < max_reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        reps.append(time.perf_counter() - t0)
    return reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dp45-steps", type=int, default=12)
    ap.add_argument("--size", default="1024x1024x4")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--heff-only", choices=["f32", "f64", "f32_full"], default=None,
                    help="run --steps H_eff evaluations of one path and exit (for profiler runs); f32_full: "
                         "the FP32 field on the reference tensor's spectrum (DemagField(full_tensor=True))")
    ap.add_argument("--dp45-only", choices=["f32", "f64", "f32_full"], default=None,
                    help="run --steps DP45 steps of one path and exit (for profiler runs)")
    ap.add_argument("--c5-precisions", default="f32,f64", help="C5: paths whose H_eff rate is measured")
    ap.add_argument("--config", choices=["C1", "C2traj", "C3", "C4", "C5", "C5full"], default=None,
                    help="run one secondary configuration and print its JSON line instead of the main one")
    ap.add_argument("--c2-seeds", default="2024", help="C2traj: comma-separated seeds of the initial state")
    ap.add_argument("--c2-paths", default="f32,f32_full,f32_demag,f32_stencils,f64_oct",
                    help="C2traj: paths to compare with the FP64 path")
    ap.add_argument("--c4-samples", type=int, default=4096, help="C4: training films (64x64x1)")
    ap.add_argument("--c4-epochs", type=int, default=100, help="C4: training epochs (cosine step size 1e-3 -> 2e-5; the best held-out epoch is kept)")
    ap.add_argument("--c4-eval", type=int, default=16, help="C4: 256x256x1 evaluation films")
    ap.add_argument("--c4-target-torque", type=float, default=1e-6, help="C4: torque of the training targets")
    ap.add_argument("--c4-weights", default=None, help="C4: write the trained parameters here (.npy); "
                    "C5full: read them from here when the file exists")
    ap.add_argument("--c5-samples", type=int, default=256, help="C5full: training films when no weights are given")
    ap.add_argument("--c5-lbfgs-iters", type=int, default=4000, help="C5full: LBFGS iteration bound")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import magnumaf as M

    if args.config:
        print(json.dumps(CONFIGS[args.config](M, torch, args)))
        return

    nx, ny, nz = (int(v) for v in args.size.split("x"))
    d = 2e-9
    Ms, A, Ku, D, alpha = 1.1e6, 1.5e-11, 8e5, 3e-3, 0.1
    mesh = M.Mesh(nx, ny, nz, d, d, d)
    m0 = workload(nx, ny, nz)
    terms = [M.DemagField(mesh), M.ExchangeField(A), M.DmiField(D, [0, 0, 1]), M.UniaxialAnisotropyField(Ku, [0, 0, 1])]

    t0 = time.time()
    state = M.State(mesh, Ms, m0)
    if args.dp45_only:
        tl = ([M.DemagField(mesh, full_tensor=True)] + terms[1:]) if args.dp45_only == "f32_full" else terms
        llg = M.LLGIntegrator(alpha, tl, "DP45", precision=args.dp45_only.replace("_full", ""))
        for _ in range(args.warmup + args.steps):
            llg.step(state)
        torch.cuda.synchronize()
        log(f"[bench] {args.steps} {args.dp45_only} DP45 steps done ({time.time() - t0:.1f} s)")
        return
    if args.heff_only:
        if args.heff_only == "f32_full":
            terms = [M.DemagField(mesh, full_tensor=True)] + terms[1:]
        ctx = state._context(terms, args.heff_only.replace("_full", ""))
        f32out = args.heff_only != "f64"  # each path writes its field in its own precision
        h = torch.empty(3 * mesh.n, dtype=torch.float32 if f32out else torch.float64, device="cuda")
        for _ in range(args.warmup + args.steps):
            ctx.heff(state.m_dev, out=h, f32=f32out)
        torch.cuda.synchronize()
        # a plain device-to-device copy of known size: the reference point for HBM byte counters
        src = torch.ones(COPY_CAL_BYTES // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] {args.steps} {args.heff_only} H_eff evaluations and 3 copies of {COPY_CAL_BYTES} B done "
            f"({time.time() - t0:.1f} s)")
        return
    ctx32 = state._context(terms, "f32")
    log(f"[bench] FP32 context ready in {time.time() - t0:.1f} s (includes host Newell tensor)")
    t0 = time.time()
    ctx64 = state._context(terms, "f64")
    log(f"[bench] FP64 context ready in {time.time() - t0:.1f} s")
    N = mesh.n
    # Each path writes H_eff in its own precision (the FP32-field path 3N float32, the FP64 path 3N
    # float64), as the reference does for f32 and f64 m (demag_field.cpp:217-221) and as the
    # algorithmic byte count of the path has it (3N s for the output).
    k32 = torch.empty(3 * N, dtype=torch.float64, device="cuda")
    h32 = torch.empty(3 * N, dtype=torch.float32, device="cuda")

    def heff_loop(ctx, n):
        f32out = ctx.precision != M._abi.PRECISIONS["f64"]
        for _ in range(n):
            ctx.heff(state.m_dev, out=h32 if f32out else k32, f32=f32out)

    # ---------------------------------------------------------------- timed region: FP32 H_eff
    # Each repeat times exactly --steps evaluations between two synchronisations; repeats continue
    # until at least MIN_TIMED_S of work has been timed (and at least 3 repeats), and the median
    # repeat sets the rate, so a short --steps still measures a steady state.
    heff_loop(ctx32, args.warmup)
    torch.cuda.synchronize()
    reps32 = timed_repeats(lambda: heff_loop(ctx32, args.steps), torch)
    t_fp32 = float(np.median(reps32))
    heff32 = args.steps / t_fp32
    log(f"[bench] FP32-field H_eff: {heff32:.1f} /s ({1e3 * t_fp32 / args.steps:.4f} ms; median of {len(reps32)} "
        f"repeats of {args.steps}, {sum(reps32):.2f} s timed)")

    # ---------------------------------------------------------------- FP64 H_eff
    k64n = max(20, args.steps // 4)
    heff_loop(ctx64, 3)
    torch.cuda.synchronize()
    reps64 = timed_repeats(lambda: heff_loop(ctx64, k64n), torch)
    heff64 = k64n / float(np.median(reps64))
    log(f"[bench] FP64 H_eff: {heff64:.1f} /s")
    # the FP32-field path with its field written as float64 (the form of the FP64 path's output)
    def heff_loop_f64out(n):
        for _ in range(n):
            ctx32.heff(state.m_dev, out=k32)
    heff_loop_f64out(3)
    torch.cuda.synchronize()
    heff32_f64out = args.steps / float(np.median(timed_repeats(lambda: heff_loop_f64out(args.steps), torch)))
    log(f"[bench] FP32-field H_eff, float64 output: {heff32_f64out:.1f} /s")

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
    # HBM bytes per launch of the dominant kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    # runs at this workload (profiles/heff_*_1024x1024x4_hbm_traffic.md), when it is the default one
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "heff_1024x1024x4_traffic.json")
    if (nx, ny, nz) == (1024, 1024, 4) and os.path.exists(tfile):
        with open(tfile) as f:
            t = json.load(f).get("f32", {}).get(dom[0])
        if t:
            traffic = t["read_bytes"] + t["write_bytes"]
    # FP32-field H_eff error against the FP64 path (normwise: max|dH| / max|H64|) on the initial state
    h64 = torch.empty_like(k32)
    ctx64.heff(state.m_dev, out=h64)
    ctx32.heff(state.m_dev, out=h32, f32=True)
    torch.cuda.synchronize()
    heff_err32 = ((h32.double() - h64).abs().max() / h64.abs().max()).item()
    del h64
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
    def dp45_rate(precision, nsteps, tl=terms):
        st = M.State(mesh, Ms, m0)
        llg = M.LLGIntegrator(alpha, tl, "DP45", precision=precision)
        st._ctx = state._ctx  # reuse the contexts built above
        for _ in range(3):
            llg.step(st)
        torch.cuda.synchronize()
        s0 = llg.stats()

        def run():
            for _ in range(nsteps):
                llg.step(st)
        reps = timed_repeats(run, torch)
        s1 = llg.stats()
        dt = sum(reps)
        attempts = (s1["heff_evals"] - s0["heff_evals"]) / 7
        return nsteps / float(np.median(reps)), attempts / dt, s1

    dp32, at32, s32 = dp45_rate("f32", args.dp45_steps)
    log(f"[bench] DP45 FP32-field: {dp32:.2f} steps/s ({at32:.2f} attempts/s)")
    dp64, at64, s64 = dp45_rate("f64", max(4, args.dp45_steps // 2))
    log(f"[bench] DP45 FP64: {dp64:.2f} steps/s ({at64:.2f} attempts/s)")

    # ---------------------------------------------------------------- FP32 field on the reference tensor's spectrum
    # MAFX_TERM_FULL_TENSOR: the FP32 arithmetic on the reference's own (complex) tensor spectrum
    terms_full = [M.DemagField(mesh, full_tensor=True)] + terms[1:]
    ctxf = state._context(terms_full, "f32")
    heff_loop(ctxf, 5)
    torch.cuda.synchronize()
    repsf = timed_repeats(lambda: heff_loop(ctxf, args.steps), torch)
    heff32f = args.steps / float(np.median(repsf))
    dp32f, at32f, _ = dp45_rate("f32", args.dp45_steps, terms_full)
    log(f"[bench] FP32 field on the full tensor spectrum: {heff32f:.1f} H_eff/s, DP45 {dp32f:.2f} steps/s")
    del state._ctx[(tuple(id(t) for t in terms_full), "f32")], ctxf

    # ---------------------------------------------------------------- CPU baseline (oracle)
    cpu = None
    sim = None
    ref_total = None
    if not args.no_cpu:
        import oracle_lib as O
        threads = int(os.environ.get("OMP_NUM_THREADS", str(os.cpu_count() or 1)))
        t0 = time.time()
        sim = O.Sim((nx, ny, nz, d, d, d), Ms, m0,
                    [O.term(O.DEMAG), O.term(O.EXCHANGE, A), O.term(O.DMI, D), O.term(O.UNIAXIAL, Ku)],
                    nthreads=threads)
        log(f"[bench] oracle setup {time.time() - t0:.1f} s")
        t0 = time.perf_counter()
        ref_total = sim.heff_flat()
        cpu_s = time.perf_counter() - t0
        cpu = {"value": 1.0 / cpu_s, "unit": "H_eff evals/s", "cores": threads, "kind": "port",
               "sample": f"one FP64 H_eff of the oracle, the unoptimised correctness restatement of the reference "
                         f"(complex 3-D FFTs through a recursive mixed-radix DFT, 27-tap stencils; not a tuned CPU "
                         f"path), at the full {nx}x{ny}x{nz} mesh on {threads} host threads; tensor setup untimed"}
        log(f"[bench] oracle H_eff {cpu_s:.2f} s")

    # ---------------------------------------------------------------- field errors per term
    errors = field_errors(M, torch, mesh, state, terms, ctx32, ctx64, sim, ref_total)

    out = {
        "metric": "H_eff evals/s & DP45 steps/s at 1024^2x4 cells; LBFGS time-to-tolerance",
        "value": heff32,
        "unit": "H_eff evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_fp32 / args.steps,
        "repeats": len(reps32),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 field / f64 state",
        "data": "synthetic (seeded random unit vectors)",
        "config": {"workload": f"C2 film {nx}x{ny}x{nz} cells of 2 nm, demag+exchange+DMI+uniaxial, alpha=0.1",
                   "cells": N, "path": "FP32-field H_eff"},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": dom[1]["GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom[1]["GBps"] / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": dom[1]["bytes"], "avg_ms": dom[1]["avg_ms"]},
        "cpu_baseline": cpu,
        "extra": {
            "heff_per_s_fp32": heff32, "heff_per_s_fp64": heff64, "heff_per_s_fp32_float64_output": heff32_f64out,
            "dp45_steps_per_s_fp32": dp32, "dp45_steps_per_s_fp64": dp64,
            "dp45_attempts_per_s_fp32": at32, "dp45_attempts_per_s_fp64": at64,
            "dp45_fp32_over_fp64": dp32 / dp64 if dp64 > 0 else None,
            "heff_per_s_fp32_full_tensor": heff32f, "dp45_steps_per_s_fp32_full_tensor": dp32f,
            "dp45_fp32_full_tensor_over_fp64": dp32f / dp64 if dp64 > 0 else None,
            "heff_fp32_normwise_error_vs_fp64": heff_err32,
            "field_errors": errors,
            "heff_algorithmic_GB_fp32": bytes_heff32 / 1e9, "heff_algorithmic_GB_fp64": bytes_heff64 / 1e9,
            "heff_kernel_ms_fp32": per_heff_ms32, "heff_kernel_ms_fp64": per_heff_ms64,
            "heff_frac_of_hbm_peak_fp32": bytes_heff32 / (per_heff_ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "heff_frac_of_hbm_peak_fp64": bytes_heff64 / (per_heff_ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            # the same bytes over the timed loop's wall clock per evaluation (kernels overlap at
            # their boundaries, so this exceeds the sum of event-timed kernel durations' figure)
            "heff_wall_frac_of_hbm_peak_fp32": bytes_heff32 * heff32 / 1e9 / HBM_PEAK_GBS,
            "heff_wall_frac_of_hbm_peak_fp64": bytes_heff64 * heff64 / 1e9 / HBM_PEAK_GBS,
            "copy_kernel_GBps": copy_gbs,
            "kernels_fp32": kt32, "kernels_fp64": kt64,
        },
    }
    print(json.dumps(out))


def field_errors(M, torch, mesh, state, terms, ctx32, ctx64, sim, ref_total):
    """Normwise errors (max|dH| / max|H_ref|) of each term's field and of H_eff at the start state,
    against the oracle's FP64 fields (or the GPU FP64 path, which matches the oracle to 1e-12, when
    the oracle is not run), and the demag error split into the tensor's far-field noise and the FP32
    arithmetic (demag_split below)."""
    names = ["demag", "exchange", "dmi", "uniaxial"]
    n3 = 3 * mesh.n
    buf = torch.empty(n3, dtype=torch.float64, device="cuda")

    def gpu(ctx, mask):
        ctx.heff(state.m_dev, mask=mask, out=buf)
        torch.cuda.synchronize()
        return buf.cpu().numpy()

    def nw(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

    ref_name = "oracle" if sim is not None else "gpu_fp64"
    out = {"reference": ref_name, "terms": {}}
    ref_demag = None
    for k, name in enumerate(names):
        mask = 1 << k
        h64 = gpu(ctx64, mask)
        ref = sim.heff_flat(mask) if sim is not None else h64
        h32 = gpu(ctx32, mask)
        out["terms"][name] = {"max_abs_H_A_per_m": float(np.max(np.abs(ref))),
                              "fp64_vs_ref": nw(h64, ref), "fp32_vs_ref": nw(h32, ref),
                              "fp32_max_abs_err_A_per_m": float(np.max(np.abs(h32 - ref)))}
        if name == "demag":
            ref_demag = ref
            h32_default = h32
        log(f"[bench] {name}: FP64 {out['terms'][name]['fp64_vs_ref']:.2e}, FP32 {out['terms'][name]['fp32_vs_ref']:.2e} "
            f"normwise vs {ref_name}")
    h64 = gpu(ctx64, 0xFFFFFFFF)
    href = ref_total if ref_total is not None else h64
    out["heff"] = {"fp64_vs_ref": nw(h64, href), "fp32_vs_ref": nw(gpu(ctx32, 0xFFFFFFFF), href)}
    # demag split.  Every FP64 evaluation of the Newell tensor carries far-field cancellation noise
    # whose bits depend on the operation order and the libm; the reference's field (the oracle's,
    # matched by the FP64 path to 1e-15) is one of them.  Other equally valid tensors: the
    # parity-extended host octant (the FP32 path's default, differing only at negative offsets)
    # and the formulas evaluated on the GPU.  The FP32 path against the FP64 pipeline on its own
    # tensor isolates the FP32 arithmetic.
    def demag_with(prec, **kw):
        c = state._context([M.DemagField(mesh, **kw)], prec)
        h = gpu(c, 1)
        del state._ctx[(tuple(id(t) for t in c.terms), prec)]
        return h

    d64oct = demag_with("f64", octant_tensor=True)
    d64dev = demag_with("f64", device_tensor=True)
    d32dev = demag_with("f32", device_tensor=True)
    d32host = demag_with("f32", host_tensor=True)
    d32full = demag_with("f32", full_tensor=True)
    out["demag_split"] = {
        "fp64_octant_tensor_vs_ref": nw(d64oct, ref_demag),
        "fp64_gpu_tensor_vs_ref": nw(d64dev, ref_demag),
        "fp32_vs_fp64_on_the_octant_tensor": nw(h32_default, d64oct),  # FP32 arithmetic alone
        "fp32_gpu_tensor_vs_fp64_gpu_tensor": nw(d32dev, d64dev),
        "fp32_octant_tensor_vs_ref": out["terms"]["demag"]["fp32_vs_ref"],
        "fp32_whole_host_tensor_vs_ref": nw(d32host, ref_demag),
        "fp32_gpu_tensor_vs_ref": nw(d32dev, ref_demag),
        "fp32_full_tensor_vs_ref": nw(d32full, ref_demag)}  # FP32 arithmetic on the reference's tensor
    log(f"[bench] demag split: {out['demag_split']}")
    return out


def config_c1(M, torch, args):
    """128x32x1 film, DP45 (rtol = atol = 1e-6), 1 ns of dense output every 10 ps on the FP64 path,
    against the oracle on the host cores (the reference's CPU-runnable configuration)."""
    import oracle_lib as O
    nx, ny, nz = 128, 32, 1
    mesh_t = (nx, ny, nz, 4e-9, 4e-9, 3e-9)
    rng = np.random.default_rng(11)
    m0 = np.zeros((nx, ny, nz, 3))
    m0[..., 0] = 1.0
    m0 += 0.1 * rng.standard_normal(m0.shape)
    m0 /= np.linalg.norm(m0, axis=-1, keepdims=True)
    h = 25e-3 / MU0
    H = (h * math.cos(math.radians(170)), h * math.sin(math.radians(170)), 0.0)
    mesh = M.Mesh(*mesh_t)
    state = M.State(mesh, 8e5, m0)
    llg = M.LLGIntegrator(0.02, [M.DemagField(mesh), M.ExchangeField(1.3e-11), M.ExternalField(homogeneous=H)], "DP45")
    llg.integrate_dense(state, 1e-11, 1e-11)  # warm-up (context, first steps)
    state = M.State(mesh, 8e5, m0)
    llg = M.LLGIntegrator(0.02, llg.terms, "DP45")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rows = llg.integrate_dense(state, 1e-9, 1e-11)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    steps = llg.stats()["accepted"]
    log(f"[C1] GPU: {len(rows)} rows, {steps} steps in {t_gpu:.2f} s")
    threads = int(os.environ.get("OMP_NUM_THREADS", str(os.cpu_count() or 1)))
    sim = O.Sim(mesh_t, 8e5, m0, [O.term(O.DEMAG), O.term(O.EXCHANGE, 1.3e-11), O.term(O.ZEEMAN, field=H)],
                nthreads=threads)
    sim.llg_init(0.02, "DP45")
    t0 = time.perf_counter()
    ref = sim.integrate_dense(1e-9, 1e-11)
    t_cpu = time.perf_counter() - t0
    dev = float(np.max(np.abs(rows[:, 1:] - ref[:, 1:]))) if rows.shape == ref.shape else None
    log(f"[C1] oracle: {len(ref)} rows in {t_cpu:.2f} s; max |<m>_gpu - <m>_oracle| = {dev}")
    return {"config": "C1 128x32x1 DP45 1 ns dense output every 10 ps, FP64 path vs oracle",
            "gpu_s": t_gpu, "gpu_steps": steps, "gpu_steps_per_s": steps / t_gpu, "rows": len(rows),
            "max_abs_mean_m_deviation_vs_oracle": dev, "oracle_s": t_cpu, "oracle_threads": threads,
            "final_mean_m_gpu": rows[-1, 1:].tolist(), "final_mean_m_oracle": ref[-1, 1:].tolist()}


def config_c2traj(M, torch, args):
    """1024x1024x4 film, DP45 for 100 ps from seeded random unit vectors on the FP64 path and on the
    FP32 paths; <m> every 1 ps through dense output; max deviation of <m> from the FP64 run of the
    same seed.  Paths: "f32" (FP32 demag and stencils), "f32_demag" (FP32 demag, FP64 stencils),
    "f32_stencils" (FP64 demag, FP32 stencils), "f32_host" (FP32-field on the whole host tensor,
    symmetrised), "f32_full" (FP32-field on the reference tensor's own complex spectrum),
    "f64_dev" / "f64_oct" (the FP64 path on another evaluation of the tensor: the floor any path
    that does not reproduce the reference's far-field rounding bit for bit can reach)."""
    nx, ny, nz, d = 1024, 1024, 4, 2e-9
    mesh = M.Mesh(nx, ny, nz, d, d, d)
    paths = args.c2_paths.split(",")
    seeds = [int(v) for v in args.c2_seeds.split(",")]
    local = [M.ExchangeField(1.5e-11), M.DmiField(3e-3, [0, 0, 1]), M.UniaxialAnisotropyField(8e5, [0, 0, 1])]
    result = {"config": "C2 1024x1024x4 DP45 100 ps, FP32 paths vs the FP64 path, <m> every 1 ps", "seeds": {}}
    for seed in seeds:
        m0 = workload(nx, ny, nz, seed=seed)
        rows, out = {}, {}
        for path in ["f64"] + [p for p in paths if p != "f64"]:
            # f64_dev / f64_oct: the FP64 path on the GPU-built tensor or the parity-extended host
            # octant, equally valid evaluations of the Newell tensor with other far-field noise
            prec = {"f32_host": "f32", "f32_full": "f32", "f64_dev": "f64", "f64_oct": "f64"}.get(path, path)
            terms = [M.DemagField(mesh, host_tensor=path == "f32_host", device_tensor=path == "f64_dev",
                                  octant_tensor=path == "f64_oct", full_tensor=path == "f32_full")] + local
            state = M.State(mesh, 1.1e6, m0)
            llg = M.LLGIntegrator(0.1, terms, "DP45", precision=prec)
            t0 = time.perf_counter()
            chunks = []
            for c in range(10):  # 10 chunks of 10 ps, dense output every 1 ps
                r = llg.integrate_dense(state, 1e-11, 1e-12)
                chunks.append(r if c == 0 else r[1:])
            torch.cuda.synchronize()
            rows[path] = np.concatenate(chunks)
            st = llg.stats()
            out[path] = {"wall_s": time.perf_counter() - t0, "steps": st["accepted"], "rejected": st["rejected"],
                         "final_mean_m": rows[path][-1, 1:].tolist()}
            if path != "f64":
                a, b = rows["f64"], rows[path]
                n = min(len(a), len(b))
                out[path]["max_abs_mean_m_deviation"] = float(np.max(np.abs(a[:n, 1:] - b[:n, 1:])))
                out[path]["deviation_at_ps"] = [float(np.max(np.abs(a[k, 1:] - b[k, 1:]))) for k in range(0, n, 10)]
            log(f"[C2traj] seed {seed} {path}: {out[path]}")
            del llg, state
        result["seeds"][str(seed)] = out
    for path in paths:
        if path == "f64":
            continue
        devs = [result["seeds"][str(sd)][path]["max_abs_mean_m_deviation"] for sd in seeds]
        result[f"max_deviation_{path}"] = max(devs)
    return result


def config_c3(M, torch, args):
    """Atomistic 512x512x1 spins, LBFGS from seeded random spins until max|m x H| / max|H| <= 1e-6."""
    nx, ny, nz, a = 512, 512, 1, 0.27e-9
    J = 1.6e-21
    mus = 2.6 * 9.274009994e-24
    mesh = M.Mesh(nx, ny, nz, a, a, a)
    m0 = workload(nx, ny, nz, seed=3)
    terms = [M.AtomisticExchangeField(J), M.AtomisticUniaxialAnisotropyField(0.02 * J, (0, 0, 1)),
             M.AtomisticDmiField(0.2 * J, (0, 0, -1)), M.AtomisticDipoleDipoleField(mesh),
             M.AtomisticExternalField(homogeneous=(0, 0, 1.0 / MU0))]
    res = {}
    for prec in ["f64"]:
        state = M.State(mesh, mus, m0)
        mini = M.LBFGS_Minimizer(terms, tol=1e-20, maxiter=200000, torque_tol=1e-6, precision=prec)
        state._context(terms, prec)  # tensor setup outside the timed region
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E = mini.minimize(state)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[prec] = dict(mini.last_stats, wall_s=dt, E=E)
        log(f"[C3] {prec}: {dt:.2f} s, {mini.last_stats}")
    return {"config": "C3 atomistic 512x512x1 LBFGS to max|m x H|/max|H| <= 1e-6 (DMI axis (0,0,-1))",
            "lbfgs": res}


def config_c5(M, torch, args):
    """2048x1024x8 multilayer (RKKY between layers), H_eff/s of both paths and DP45 steps/s (FP32 field)."""
    nx, ny, nz, d = 2048, 1024, 8, 2e-9
    mesh = M.Mesh(nx, ny, nz, d, d, d)
    m0 = workload(nx, ny, nz, seed=5)
    rk = np.full((nx, ny, nz), -1e-3 * d / 2.0)
    ex = np.full((nx, ny, nz), 1.3e-11)
    terms = [M.DemagField(mesh), M.RKKYExchangeField(rk, ex, mesh), M.DmiField(2.5e-3, [0, 0, 1]),
             M.UniaxialAnisotropyField(7e5, [0, 0, 1]), M.ExternalField(homogeneous=(50e-3 / MU0, 0, 0))]
    state = M.State(mesh, 1.0e6, m0)
    out = {}
    for prec in args.c5_precisions.split(","):
        t0 = time.time()
        ctx = state._context(terms, prec)
        log(f"[C5] {prec} context in {time.time() - t0:.1f} s")
        h = torch.empty(3 * mesh.n, dtype=torch.float64, device="cuda")
        for _ in range(3):
            ctx.heff(state.m_dev, out=h)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            ctx.heff(state.m_dev, out=h)
        torch.cuda.synchronize()
        out[f"heff_per_s_{prec}"] = 20 / (time.perf_counter() - t0)
        log(f"[C5] {prec} H_eff/s {out[f'heff_per_s_{prec}']:.1f}")
    st = M.State(mesh, 1.0e6, m0)
    st._ctx = state._ctx
    llg = M.LLGIntegrator(0.05, terms, "DP45", precision="f32")
    for _ in range(2):
        llg.step(st)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(6):
        llg.step(st)
    torch.cuda.synchronize()
    out["dp45_steps_per_s_fp32"] = 6 / (time.perf_counter() - t0)
    for key, c in state._ctx.items():
        out[f"device_bytes_{key[1]}_context"] = c.device_bytes()
    log(f"[C5] DP45 FP32-field steps/s {out['dp45_steps_per_s_fp32']:.2f}")
    return {"config": "C5 2048x1024x8 RKKY multilayer, demag+RKKY+DMI+uniaxial+50 mT", **out}


def config_c5full(M, torch, args):
    """C5 in its full form: a learned warm start of the 2048x1024x8 RKKY multilayer (network trained
    as in C4 on a small set of relaxed 64x64x1 films, applied plane by plane), LBFGS on the FP32-field
    path to max|m x H| / max|H| <= 1e-6 (bounded iterations), then 50 ps of DP45 (FP32 field, FP64
    state, alpha = 0.05) under the 50 mT in-plane field."""
    import magnumaf.warmstart as WS
    nx, ny, nz, d = 2048, 1024, 8, 2e-9
    out = {}
    t0 = time.perf_counter()
    if args.c4_weights and os.path.exists(args.c4_weights):
        net = WS.WarmStartNet.load(args.c4_weights)
        out["network"] = f"loaded from {args.c4_weights}"
    else:
        data = WS.generate_dataset(args.c5_samples, seed=1, size=64, torque_tol=1e-4, log=log)
        net = WS.WarmStartNet()
        WS.train(net, data, epochs=args.c4_epochs, batch=32, lr=1e-3, log=log)
        out["network"] = f"trained on {args.c5_samples} films of 64x64x1, {args.c4_epochs} epochs"
    out["network_s"] = time.perf_counter() - t0
    log(f"[C5full] network ready in {out['network_s']:.1f} s ({out['network']})")
    mesh = M.Mesh(nx, ny, nz, d, d, d)
    m0 = workload(nx, ny, nz, seed=5)
    rk = np.full((nx, ny, nz), -1e-3 * d / 2.0)
    ex = np.full((nx, ny, nz), 1.3e-11)
    terms = [M.DemagField(mesh), M.RKKYExchangeField(rk, ex, mesh), M.DmiField(2.5e-3, [0, 0, 1]),
             M.UniaxialAnisotropyField(7e5, [0, 0, 1]), M.ExternalField(homogeneous=(50e-3 / MU0, 0, 0))]
    state = M.State(mesh, 1.0e6, m0)
    t0 = time.perf_counter()
    state._context(terms, "f32")
    out["context_s"] = time.perf_counter() - t0
    # the FP64 context the FP32 minimisation finishes on (its reference-order host tensor is
    # setup, like the FP32 one): built here, so that the LBFGS timing holds the minimisation only
    t0 = time.perf_counter()
    state._context(terms, "f64")
    out["finish_context_s"] = time.perf_counter() - t0
    log(f"[C5full] contexts: FP32 {out['context_s']:.1f} s, FP64 finish {out['finish_context_s']:.1f} s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    WS.warm_start_layers(net, state, {"Ms": 1.0e6, "A": 1.3e-11, "D": 2.5e-3, "Ku": 7e5, "Bz": 0.0})
    torch.cuda.synchronize()
    out["warm_start_s"] = time.perf_counter() - t0
    mini = M.LBFGS_Minimizer(terms, tol=1e-20, maxiter=args.c5_lbfgs_iters, verbose=2, torque_tol=1e-6,
                             precision="f32")
    t0 = time.perf_counter()
    E = mini.minimize(state)
    torch.cuda.synchronize()
    out["lbfgs_s"] = time.perf_counter() - t0
    st = mini.last_stats
    out.update(lbfgs_iterations=st["iterations"], lbfgs_evals=st["energy_evals"], lbfgs_stop_reason=st["stop_reason"],
               lbfgs_final_torque_rel=st["final_torque_rel"], lbfgs_energy=E)
    log(f"[C5full] LBFGS: {st['iterations']} iterations in {out['lbfgs_s']:.1f} s, torque {st['final_torque_rel']:.3e}, "
        f"stop reason {st['stop_reason']}")
    llg = M.LLGIntegrator(0.05, terms, "DP45", precision="f32")
    t_start = state.t
    t0 = time.perf_counter()
    steps = 0
    for k in range(1, 11):
        steps += llg.run_until(state, t_start + k * 5e-12)
        log(f"[C5full] DP45 t = {state.t - t_start:.3e} s, {steps} steps, <m> = {state.mean_m()}")
    torch.cuda.synchronize()
    out["dp45_s"] = time.perf_counter() - t0
    out.update(dp45_steps=steps, dp45_simulated_s=state.t - t_start, dp45_steps_per_s=steps / out["dp45_s"],
               mean_m_final=[float(v) for v in state.mean_m()], rejected=llg.stats()["rejected"])
    out["end_to_end_s"] = out["warm_start_s"] + out["lbfgs_s"] + out["dp45_s"]
    return {"config": "C5 full form: warm start + LBFGS + 50 ps DP45, 2048x1024x8 RKKY multilayer", **out}


def config_c4(M, torch, args):
    """Learned warm start: FP64 LBFGS relaxations of seeded 64x64x1 films as targets, training of
    the encoder-decoder on the GPU, then warm start + LBFGS vs plain LBFGS at 256x256x1."""
    import magnumaf.warmstart as WS
    t0 = time.perf_counter()
    data = WS.generate_dataset(args.c4_samples, seed=1, size=64, torque_tol=args.c4_target_torque, log=log)
    t_data = time.perf_counter() - t0
    its = [mm["iterations"] for mm in data["meta"]]
    log(f"[C4] dataset: {args.c4_samples} films in {t_data:.1f} s, LBFGS iterations median {np.median(its):.0f}")
    hold_idx, _ = WS.split(len(its))
    base = WS.baseline_losses(data, hold_idx)
    log(f"[C4] held-out loss of predictors without training: {base}")
    net = WS.WarmStartNet()
    t0 = time.perf_counter()
    hist, hold = WS.train(net, data, epochs=args.c4_epochs, batch=32, lr=1e-3, lr_min=2e-5, log=log)
    t_train = time.perf_counter() - t0
    if args.c4_weights:
        net.save(args.c4_weights)
    ev = WS.evaluate(net, n=args.c4_eval, size=256, log=log, perturb=1e-6)
    sp = [e["speedup"] for e in ev]
    same = [e["E_rel_diff"] <= 1e-9 for e in ev]
    same_pert = [e["E_rel_diff_perturbed"] <= 1e-9 for e in ev]
    return {"config": "C4 learned warm start: 64x64x1 training films, 256x256x1 evaluation, torque <= 1e-6",
            "samples": args.c4_samples, "target_torque": args.c4_target_torque, "dataset_s": t_data,
            "train_s": t_train, "epochs": args.c4_epochs, "train_loss": hist, "heldout_loss": hold,
            "heldout_loss_without_training": base, "eval": ev,
            "speedup_median": float(np.median(sp)), "speedup_geomean": float(np.exp(np.mean(np.log(sp)))),
            "equal_final_energy_count": int(sum(same)),
            "plain_from_perturbed_start_equal_energy_count": int(sum(same_pert)),
            "warm_lower_energy_count": int(sum(e["E_warm"] < e["E_plain"] and not s for e, s in zip(ev, same))),
            "lbfgs_iterations_median_dataset": float(np.median(its))}


CONFIGS = {"C1": config_c1, "C2traj": config_c2traj, "C3": config_c3, "C4": config_c4, "C5": config_c5,
           "C5full": config_c5full}


if __name__ == "__main__":
    main()
