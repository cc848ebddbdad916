"""Builds the gfx950 library tosca_amd/libtosca_rhs.so with hipcc (no torch extension machinery).

    python -m tosca_amd.build            # or tosca_amd.build.build()

Each .hip source is compiled to an object in tosca_amd/_build/ and linked into one shared
library next to this file, so it ships with the source tree.  The reference-structured kernels
are compiled with FP contraction off (they track the CPU oracle bit for bit); the fused kernels
use the compiler's default contraction.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libtosca_rhs.so"
ARCH = os.environ.get("TOSCA_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-variable", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
]

SOURCES = {
    "api.hip": [],
    "ref_kernels.hip": ["-ffp-contract=off"],
    "fused_formu.hip": [],
    "fused_les.hip": [],
    "vjp.hip": [],
    "explicit_rhs.hip": ["-ffp-contract=off"],
    "bc_fill.hip": ["-ffp-contract=off"],
    "formt.hip": ["-ffp-contract=off"],
    "gradp.hip": ["-ffp-contract=off"],
    "krylov.hip": ["-ffp-contract=off"],
    "farm.hip": ["-ffp-contract=off"],
    "rk3w.hip": ["-ffp-contract=off"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def build(verbose: bool = False) -> Path:
    cc = hipcc()
    BUILD.mkdir(exist_ok=True)
    objs = []
    for src, extra in SOURCES.items():
        obj = BUILD / (Path(src).stem + ".o")
        cmd = [cc, *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
    print(LIB)
