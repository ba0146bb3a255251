"""Build libtflows_krylov.so for gfx950 in-tree with hipcc.

Every source is compiled for MI355X only.  Flags:
* ``-ffp-contract=off`` -- no multiply-add is fused, so element-wise updates
  and the row sums of the SpMV round exactly like the reference Fortran;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops`` -- packed FP32
  instructions are kept out of the generated code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(REPO, "_build")
LIB = os.path.join(PKG, "libtflows_krylov.so")
FORTRAN = os.path.join(PKG, "fortran")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
ARCH = "gfx950"
SOURCES = ["krylov_kernels.hip", "cg_fused.hip", "system.hip", "reorder.hip", "capi.hip", "mixed.hip", "batched.hip",
           "adjoint.hip", "calcps.hip", "meshio.cpp", "noalloc.hip", "tile_plan.cpp", "slot_layout.cpp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I", os.path.join(REPO, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the ISO_C_BINDING interface module and the drop-in driver
    (a Fortran program calling cg/bicg/cgs the way CalcPS.f90:217 and
    NewUVW.f90:506 do) against the library, if amdflang is present."""
    if not os.path.exists(FLANG) or not os.path.isdir(FORTRAN):
        return None
    out = os.path.join(PKG, "tflows_dropin_driver")
    mod_dir = os.path.join(BUILD, "fmod")
    os.makedirs(mod_dir, exist_ok=True)
    srcs = [os.path.join(FORTRAN, f) for f in ("tflows_krylov_mod.f90", "dropin_explicit.f90", "dropin_driver.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    cmd = [FLANG, "-O2", "-fdefault-real-8", "-fdefault-double-8", "-module-dir", mod_dir,
           "-o", out, *srcs, "-L" + PKG, "-ltflows_krylov", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    return out


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
