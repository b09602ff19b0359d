"""Build libfvm2d.so (HIP, gfx950) in-tree.

    python 2d_fvm_cuda_amd/build.py            # build
    python 2d_fvm_cuda_amd/build.py --report   # also print kernel resource usage
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libfvm2d.so")
PROBES_LIB = os.path.join(PKG, "libfvm2d_probes.so")

# the solver library, and the measurement probes kept in a library of their own
SOURCES = ["fvm2d_step.hip", "fvm2d_adjoint.hip", "fvm2d_runtime.hip", "fvm2d_adj_runtime.hip",
           "fvm2d_noalloc.hip"]
PROBE_SOURCES = ["fvm2d_probes.hip"]
HEADERS = ["fvm2d_cell.h", "fvm2d_device.h", "fvm2d_kernels.h", "fvm2d_adjoint_cell.h",
           "fvm2d_march.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # every floating-point operation rounds once, in source order; the fast
    # path writes its fused multiply-adds explicitly with fma()
    "-ffp-contract=off",
    "-fno-fast-math",
    # packed FP32 instructions stay off (gfx950)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str, bdir: str, report: bool, verbose: bool) -> str:
    obj = os.path.join(bdir, os.path.basename(src) + ".o")
    cmd = [hipcc()] + HIPCC_FLAGS + ["-c", src, "-o", obj]
    if report:
        cmd += ["-Rpass-analysis=kernel-resource-usage"]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    if report:
        sys.stdout.write(r.stderr)
    return obj


def _link(objs, out: str) -> None:
    tmp = out + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"link of {os.path.basename(out)} failed")
    os.replace(tmp, out)


def build(report: bool = False, verbose: bool = False) -> str:
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    for srcs, out in ((SOURCES, LIB), (PROBE_SOURCES, PROBES_LIB)):
        objs = [_compile(os.path.join(CSRC, s), bdir, report, verbose) for s in srcs]
        _link(objs, out)
    return LIB


if __name__ == "__main__":
    build(report="--report" in sys.argv, verbose="-v" in sys.argv)
    print(LIB)
