"""Builds libwgc.so (HIP kernels + C ABI runtime) for gfx950 in-tree.

The kernels must reproduce the reference's binary32 arithmetic exactly, so
they are compiled without floating-point contraction (no FMA), with
correctly rounded division and without fast-math.  Packed FP32 instructions
are disabled (``-packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent
CSRC = ROOT / "csrc"
LIB = PKG_DIR / "libwgc.so"
INCLUDE = ROOT.parent / "include"

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

SOURCES = ["kernels.hip", "runtime.cpp"]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; the HIP toolchain (ROCm) is required")
    return exe


def build_library(verbose: bool = False) -> Path:
    """Compiles every source for gfx950 and links libwgc.so next to this file."""
    objs = []
    build_dir = PKG_DIR / "_build"
    build_dir.mkdir(exist_ok=True)
    for src in SOURCES:
        obj = build_dir / (src + ".o")
        cmd = [hipcc(), *HIP_FLAGS, "-x", "hip", "-I", str(INCLUDE), "-c", str(CSRC / src), "-o", str(obj)]
        run(cmd, verbose)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    run(cmd, verbose)
    os.replace(tmp, LIB)
    return LIB


def run(cmd, verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(
        line for line in (res.stdout + res.stderr).splitlines()
        if "packed-fp32-ops" not in line and line.strip()
    )
    if res.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out:
        print(out, flush=True)


if __name__ == "__main__":
    print(build_library(verbose=True))
