"""The C-ABI library builds for gfx950, loads, and exports every entry point
declared in include/tflows_krylov.h (no compute calls: CPU only)."""
import ctypes
import os
import subprocess



def test_library_exports_declared_symbols(pkg):
    k = pkg.krylov
    assert os.path.exists(k.LIB_PATH), "run __graft_entry__.build() first"
    declared = k.declared_symbols()
    assert {"cg_", "bicg_", "cgs_", "prec1_", "resid_", "tflows_solve"} <= set(declared)
    out = subprocess.run(["nm", "-D", "--defined-only", k.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(k.LIB_PATH)
    for s in declared:
        assert hasattr(lib, s)


def test_dropin_symbols_use_fortran_mangling(pkg):
    # external Fortran subroutines cg/bicg/CGS/Prec1/Resid mangle to the
    # lowercase name with one trailing underscore (no module prefix)
    out = subprocess.run(["nm", "-D", "--defined-only", pkg.krylov.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for s in ("cg_", "bicg_", "cgs_", "prec1_", "resid_"):
        assert s in names


def test_code_object_targets_gfx950(pkg):
    # the offload bundle embedded in .hip_fatbin names its target triple
    data = open(pkg.krylov.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
