"""Builds the HIP library libdazi_dk.so for gfx950 in-tree with hipcc.

Every call compiles (the driver's "does it build" check relies on that);
`ensure_built()` builds only when the library is missing.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdazi_dk.so")
SOURCES = ["dk_kernels.hip", "dk_ti_kernels.hip", "dk_grows.hip", "dk_azimuth.hip", "dk_travel.hip",
           "dk_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DAZI_DK_ARCH", "gfx950")

# -packed-fp32-ops: no packed FP32 instructions anywhere in this library.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-lambda-capture",
]


def build(verbose: bool = False) -> str:
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_hostcheck(verbose)
    return LIB


HOSTCHECK = os.path.join(HERE, "libdazi_dk_hostcheck.so")


def build_hostcheck(verbose: bool = False) -> str:
    """Host-only build of the root search's secular function (csrc/dk_fdm.h)
    for the CPU tests that pin it against the oracle; not loaded by the
    library or its bindings."""
    tmp = HOSTCHECK + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", os.path.join(CSRC, "dk_hostcheck.cpp"),
           "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, HOSTCHECK)
    return HOSTCHECK


FORTRAN = os.path.join(HERE, "fortran")
FDRIVER = os.path.join(FORTRAN, "dk_fdriver")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def build_fortran(verbose: bool = False) -> str:
    """Compiles the Fortran binding (module dazi_dk) and its example host
    dk_fdriver against libdazi_dk.so (rpath $ORIGIN/..)."""
    tmp = os.path.join(FORTRAN, "_fbuild")
    os.makedirs(tmp, exist_ok=True)
    cmds = [
        [FLANG, "-O2", "-module-dir", tmp, "-c", os.path.join(FORTRAN, "dazi_dk.f90"),
         "-o", os.path.join(tmp, "dazi_dk.o")],
        [FLANG, "-O2", "-module-dir", tmp, "-I", tmp, os.path.join(FORTRAN, "dk_fdriver.f90"),
         os.path.join(tmp, "dazi_dk.o"), "-L", HERE, "-ldazi_dk", "-Wl,-rpath,$ORIGIN/..",
         "-o", FDRIVER + ".tmp"],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(FDRIVER + ".tmp", FDRIVER)
    return FDRIVER


def ensure_built() -> str:
    if not os.path.exists(LIB):
        build()
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
