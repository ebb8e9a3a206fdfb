"""Build the HIP library libples_mg.so for gfx950 (in-tree)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libples_mg.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


FC = os.environ.get("PLES_FC", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FDRIVER = os.path.join(LIBDIR, "ples_mg_driver")


def build_fortran(verbose: bool = False) -> str:
    """Fortran binding module + the test driver, linked against libples_mg.so."""
    mod = os.path.join(LIBDIR, "fmod")
    os.makedirs(mod, exist_ok=True)
    cmds = [
        [FC, "-O2", "-c", os.path.join(FDIR, "ples_mg_f.f90"), "-J", mod, "-o", os.path.join(mod, "ples_mg_f.o")],
        [FC, "-O2", os.path.join(FDIR, "ples_mg_driver.f90"), os.path.join(mod, "ples_mg_f.o"), "-J", mod,
         "-I", mod, "-L", LIBDIR, "-lples_mg", "-Wl,-rpath," + LIBDIR, "-o", FDRIVER],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return FDRIVER


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, "ples_mg.hip"), os.path.join(CSRC, "restart.cpp")]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if os.path.exists(FC):
        build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
