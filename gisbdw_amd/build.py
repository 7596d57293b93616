"""Builds libtdhip.so (the HIP kernels and the C ABI) in-tree for gfx950.

Usage: python -m gisbdw_amd.build   (or gisbdw_amd.build.build_library())
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libtdhip.so"
INCLUDE = HERE.parent / "include"

ARCH = os.environ.get("TDHIP_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
]


def sources():
    return sorted(str(p) for p in CSRC.glob("*.hip"))


CLI = HERE / "examples" / "solve_cli"


def build_library(verbose: bool = True) -> Path:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, *HIPCC_FLAGS, *sources(), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


def build_cli(verbose: bool = True) -> Path:
    """The C++ example caller (examples/solve_cli.cpp): plain g++ against
    include/tdhip.hpp, linked to libtdhip.so with an rpath to this folder."""
    cmd = ["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", str(CLI) + ".cpp", "-o", str(CLI),
           f"-L{HERE}", "-ltdhip", f"-Wl,-rpath,{HERE}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return CLI


if __name__ == "__main__":
    build_library()
    sys.exit(0)
