"""Build the native parts of the package in-tree.

* ``lib2dense_amd.so`` -- the HIP library behind include/dense_capi.h, built
  for gfx950 with FMA contraction on (IEEE semantics otherwise: no fast-math,
  so NaN and infinity checks stay) and packed-FP32 instructions disabled.
* ``bin/dense_gpu`` -- the Fortran driver (iso_c_binding over the C ABI),
  built with AMD flang when it is installed.

Run ``python 2dense_amd/build.py`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "lib2dense_amd.so")
BIN_DIR = os.path.join(HERE, "bin")
OBJ_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FC", "/opt/rocm/lib/llvm/bin/flang")
ARCH = os.environ.get("DENSE_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["dense_kernels.hip", "dense_batch.hip", "dense_capi.hip", "dense_adjoint.hip", "dense_probe.hip"]
HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-ffp-contract=on", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    return r


def build_hip(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, src.replace(".hip", ".o"))
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    srcs = [os.path.join(FSRC, f) for f in ("dense_capi_mod.f90", "dense_driver.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found; Fortran driver not built")
        return None
    os.makedirs(BIN_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    out = os.path.join(BIN_DIR, "dense_gpu")
    cmd = [FLANG, "-O2", "-J", OBJ_DIR, *srcs, "-o", out,
           f"-L{HERE}", "-l2dense_amd", f"-Wl,-rpath,$ORIGIN/..",
           "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    return out


def build_all(verbose: bool = False):
    lib = build_hip(verbose)
    drv = build_fortran(verbose)
    return lib, drv


if __name__ == "__main__":
    lib, drv = build_all(verbose="-v" in sys.argv)
    print("built", lib, drv or "(no Fortran driver)")
