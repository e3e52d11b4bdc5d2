"""Build the HIP library in-tree for gfx950 (hipcc cross-compiles without a GPU), and the
Fortran host binding with its demo driver when a Fortran compiler is present."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "trans_hip.hip")
OUT = os.path.join(HERE, "libtrans_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay off (they have produced silent wrong results on these parts)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no FMA contraction: every kernel then performs the oracle's IEEE operation sequence and
    # the FP64 residual path is bit-identical to it
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose: bool = True) -> str:
    cmd = [HIPCC] + FLAGS + ["-o", OUT + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN_DIR, "_build")
FORTRAN_DEMO = os.path.join(FORTRAN_DIR, "_build", "box_demo")
FC = os.environ.get("FC", "/opt/rocm/lib/llvm/bin/flang")


def build_fortran(verbose: bool = True) -> str | None:
    """Compile trans_hip_mod.f90 and box_demo.f90 and link the demo against the library
    (rpath to trans_amd/).  Returns the demo's path, or None without a Fortran compiler."""
    if not os.path.exists(FC):
        if verbose:
            print(f"no Fortran compiler at {FC}: Fortran binding not built", flush=True)
        return None
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    cmd = [FC, "-O2", "-module-dir", FORTRAN_BUILD, "-o", FORTRAN_DEMO + ".tmp",
           os.path.join(FORTRAN_DIR, "trans_hip_mod.f90"), os.path.join(FORTRAN_DIR, "box_demo.f90"),
           "-L" + HERE, "-ltrans_hip", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=FORTRAN_BUILD)
    os.replace(FORTRAN_DEMO + ".tmp", FORTRAN_DEMO)
    return FORTRAN_DEMO


if __name__ == "__main__":
    build()
    build_fortran()
    sys.exit(0)
