"""Builds libmpk.so for gfx950 with hipcc, in-tree.

    python mojo-petsc4py_amd/csrc/build.py [--jobs N]

Each source compiles to an object under csrc/_build/ (in parallel), then one
link step writes mojo-petsc4py_amd/mpk/libmpk.so.  hipcc cross-compiles, so
this runs without a GPU.
"""

from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "mpk", "libmpk.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPK_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "mpk_runtime.hip",
    "mpk_vec.hip",
    "mpk_mat.hip",
    "mpk_pc.hip",
    "mpk_plan.hip",
    "mpk_assemble.hip",
    "mpk_ksp.hip",
    "mpk_mixed.hip",
    "mpk_batch.hip",
    "mpk_factor.cpp",
]

# -ffp-contract=off: every fused multiply-add in the library is an explicit
# fma(); nothing else may be contracted, so results match the CPU bit for bit.
# Packed FP32 instructions are switched off for all device code.
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-result"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def compile_one(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(HERE, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, *DEVICE, *COMMON, "-x", "hip", "-c", path, "-o", obj]
    else:
        cmd = ["g++", *COMMON, "-c", path, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stderr[-6000:]}")
    return obj


def build(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{proc.stderr[-6000:]}")
    os.replace(tmp, OUT)
    return OUT


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args()
    print(build(args.jobs))
    return 0


if __name__ == "__main__":
    sys.exit(main())
