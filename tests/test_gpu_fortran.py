"""A Fortran host calls the GPU solvers through the reference's subroutine
signatures (implicit interfaces as in CalcPS.f90:217 / NewUVW.f90:506, and
the ISO_C_BINDING module tflows_krylov): the executable links
libtflows_krylov.so in place of the reference's solver objects."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def test_fortran_dropin_driver(pkg):
    exe = os.path.join(os.path.dirname(pkg.krylov.LIB_PATH), "tflows_dropin_driver")
    if not os.path.exists(exe):
        pytest.skip("amdflang was not available when the library was built")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    lines = [l.split() for l in out.stdout.splitlines() if l.startswith("DROPIN")]
    assert [l[1] for l in lines] == ["cg", "bicg", "cgs"]
    for _, name, it, ini, fin, true_r, tol in lines:
        it, ini, fin, true_r, tol = int(it), float(ini), float(fin), float(true_r), float(tol)
        assert 0 < it < 800, name
        assert fin < tol and true_r < 1.01 * tol, name
