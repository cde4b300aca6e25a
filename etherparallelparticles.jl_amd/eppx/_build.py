"""Builds libeppx.so (the HIP library behind include/eppx.h) in-tree with hipcc for gfx950.

The pair-kernel templates are instantiated in their own translation units
(search.hip and momentum.hip, once per math mode and smoothing kernel), so
the units compile in parallel and are then linked into one shared library.
"""
from __future__ import annotations

import os
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                      # etherparallelparticles.jl_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG, "libeppx.so")

# (source, extra defines): one object each
UNITS = [("api.hip", []), ("kernels.hip", []), ("api3d.hip", []), ("kernels3d.hip", [])] + [
    (src, [f"-DEPPX_TU_FAST={fast}", f"-DEPPX_TU_KERNEL={kernel}"])
    for src in ("search.hip", "momentum.hip") for fast in (1, 0) for kernel in (1, 0)
]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # every float expression is one IEEE operation unless the code asks for fmaf
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}): {' '.join(cmd[-3:])}\n{proc.stderr[-6000:]}")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every HIP translation unit and link libeppx.so (always recompiles)."""
    jobs = jobs or min(len(UNITS), max(1, len(os.sched_getaffinity(0))))
    with tempfile.TemporaryDirectory(prefix="eppx_build_") as tmpdir:
        objs, cmds = [], []
        for k, (src, defs) in enumerate(UNITS):
            obj = os.path.join(tmpdir, f"{k:02d}_{os.path.splitext(src)[0]}.o")
            objs.append(obj)
            cmds.append([hipcc(), *HIPCC_FLAGS, *defs, "-I", INCLUDE, "-I", CSRC, "-c",
                         os.path.join(CSRC, src), "-o", obj])
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            for f in [pool.submit(_run, c, verbose) for c in cmds]:
                f.result()
        tmp = LIB + ".tmp"
        _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
