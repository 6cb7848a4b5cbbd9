"""Builds libmafx.so (HIP kernels + C-ABI) for gfx950 in-tree: magnum.af_amd/lib/libmafx.so.

    python magnum.af_amd/build.py

The host Newell tensor (csrc/newell_host.cpp) is compiled with g++ -ffp-contract=off so that its
rounding equals the reference's plain -O3 x86-64 build; every other source goes through hipcc.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libmafx.so")
ARCH = os.environ.get("MAFX_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-result", "-ffp-contract=fast-honor-pragmas"]
GXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-pthread"]

HIP_SOURCES = ["fft.hip", "field.hip", "vec.hip", "net.hip", "llg.cpp", "lbfgs.cpp", "capi.cpp"]
GXX_SOURCES = ["newell_host.cpp", "tensor_cache.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1] if cmd else ''}")
    return r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, s + ".o")
        lang = ["-x", "hip"] if s.endswith(".cpp") else []
        jobs.append((obj, [HIPCC] + HIP_FLAGS + lang + ["-c", src, "-o", obj]))
    for s in GXX_SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, s + ".o")
        jobs.append((obj, ["g++"] + GXX_FLAGS + ["-c", src, "-o", obj]))
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(workers) as ex:
        logs = list(ex.map(lambda j: _run(j[1]), jobs))
    if verbose:
        for log in logs:
            if log.strip():
                sys.stderr.write(log)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + [j[0] for j in jobs] + ["-pthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
