"""tflows_solve_noalloc against tflows_solve, through the C ABI.

tflows_solve_noalloc runs the CG kernels of tflows_solve on an existing
system with its whole workspace in the library's fixed device array, so it
allocates nothing.  It must return what tflows_solve returns, bit for bit:
x, r1, niter, IniRes and FinRes.  Two calls in a row on one process, the
second on a smaller system (the array keeps the first call's contents), in
natural and locality order, CG and BiCG, prec 0 and 1.  Sizes beyond the
maximum and the forms it does not hold are refused with an error.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


def _solve(pkg, ds, entry, method, prec, b, cap, tol, cuda):
    import torch
    kr = pkg.krylov
    lib = kr.load_library()
    f = getattr(lib, entry)
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.c_double,
                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p]
    x = torch.zeros(ds.n, dtype=torch.float64, device=cuda)
    r = torch.from_numpy(b.copy()).to(cuda)
    it = C.c_int32(cap)
    ini, fin = C.c_double(-1.0), C.c_double(-1.0)
    rc = f(ds.handle, kr.METHODS[method], prec, C.c_void_p(x.data_ptr()), C.c_void_p(r.data_ptr()), C.byref(it),
           tol, C.byref(ini), C.byref(fin), kr._stream())
    if rc != 0:
        raise kr.TFlowsError(lib.tflows_last_error().decode())
    torch.cuda.synchronize()
    return it.value, ini.value, fin.value, x.cpu().numpy(), r.cpu().numpy()


def _same(a, b):
    assert a[:3] == b[:3], (a[:3], b[:3])
    assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])


@pytest.mark.parametrize("method", ["cg", "bicg"])
@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_solve(pkg, cuda, method, prec):
    mg = pkg.meshgen
    big = mg.pressure_system(40, seed=81)                      # 64,000 rows, natural order
    small = mg.pressure_system(14, seed=82, perm_seed=5)       # 2,744 rows, stored renumbered
    cap = 3000
    for s in (big, small):
        ds = pkg.krylov.DeviceSystem(s.matrix)
        want = _solve(pkg, ds, "tflows_solve", method, prec, s.b, cap, s.tol, cuda)
        got = _solve(pkg, ds, "tflows_solve_noalloc", method, prec, s.b, cap, s.tol, cuda)
        print(f"{method} prec={prec} N={ds.n} reordered={ds.reordered} niter={got[0]} FinRes={got[2]:.3e}",
              flush=True)
        assert got[2] < s.tol
        _same(got, want)
        # a capped solve stops at cap + 1 the same way
        _same(_solve(pkg, ds, "tflows_solve_noalloc", method, prec, s.b, 9, s.tol, cuda),
              _solve(pkg, ds, "tflows_solve", method, prec, s.b, 9, s.tol, cuda))
        ds.close()


def test_noalloc_refusals(pkg, cuda):
    mg = pkg.meshgen
    s = mg.pressure_system(10, seed=83)
    ds = pkg.krylov.DeviceSystem(s.matrix)
    with pytest.raises(pkg.krylov.TFlowsError, match="CG .0. and BiCG"):
        _solve(pkg, ds, "tflows_solve_noalloc", "cgs", 1, s.b, 100, s.tol, cuda)
    with pytest.raises(pkg.krylov.TFlowsError, match="prec 0"):
        _solve(pkg, ds, "tflows_solve_noalloc", "cg", 2, s.b, 100, s.tol, cuda)
    ds.close()
    big = mg.pressure_system((MAX_ROWS // 1024 + 1, 32, 32), seed=84)
    assert big.matrix.n > MAX_ROWS
    ds = pkg.krylov.DeviceSystem(big.matrix)
    with pytest.raises(pkg.krylov.TFlowsError, match="exceeds the no-allocation maximum"):
        _solve(pkg, ds, "tflows_solve_noalloc", "cg", 1, big.b, 10, big.tol, cuda)
    ds.close()
    # the entry point still works after the refusals
    s2 = mg.pressure_system(9, seed=85)
    ds = pkg.krylov.DeviceSystem(s2.matrix)
    _same(_solve(pkg, ds, "tflows_solve_noalloc", "cg", 1, s2.b, 500, s2.tol, cuda),
          _solve(pkg, ds, "tflows_solve", "cg", 1, s2.b, 500, s2.tol, cuda))
    ds.close()
