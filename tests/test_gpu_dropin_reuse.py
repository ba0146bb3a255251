"""The drop-ins keep the device pattern between calls.

The reference fixes the matrix pattern once (TopolM, Process/Processor.f90:100)
and calls cg / bicg on it with new values every outer iteration
(CalcPS.f90:217, NewUVW.f90:506, CalcZmix.f90:394).  The drop-ins therefore
keep one device copy of the last pattern: a call on the same N, NONZ, Acol,
Arow and Adia uploads only the values and vectors.  These tests check that
such calls still match the oracle on their own values, that they give the
bits a freshly built copy gives, that prec1_ and resid_ share the copy, and
that a changed pattern is detected and rebuilt.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _inv_norm2(m):
    A = m.to_scipy().toarray()
    return 1.0 / np.linalg.svd(A, compute_uv=False)[-1]


def _check_solve(m, b, tol, got, ref):
    """The drop-in bar: niter within one, the same stopping state, x within
    1e-10 of the oracle's or within ||A^-1|| (||r_gpu|| + ||r_ref||)."""
    assert abs(got["niter"] - ref["niter"]) <= 1
    assert (got["fin_res"] < tol) == (ref["fin_res"] < tol)
    rel = _rel(got["x"], ref["x"])
    if rel > 1e-10:
        A = m.to_scipy()
        bound = _inv_norm2(m) * (np.linalg.norm(b - A @ got["x"]) + np.linalg.norm(b - A @ ref["x"]))
        assert ref["fin_res"] < tol and rel <= bound / np.linalg.norm(ref["x"]), (rel, bound)


@pytest.mark.parametrize("prec", [1, 2])
def test_same_pattern_new_values(pkg, oracle, prec):
    kr = pkg.krylov
    s = pkg.meshgen.pressure_system(12, seed=71, prec=prec)
    m1 = s.matrix
    m2 = pkg.meshgen.rescaled_pressure_values(m1, 72)
    b2 = np.random.default_rng(73).uniform(-1, 1, m1.n)
    b2 -= b2.mean()
    tol2 = 1e-8 * np.linalg.norm(b2)
    kr.dropin_release()
    st0 = kr.dropin_stats()
    first = kr.solve_system("cg", m1, s.b, prec=prec, niter=800, tol=s.tol)
    st1 = kr.dropin_stats()
    assert st1["pattern_builds"] == st0["pattern_builds"] + 1
    _check_solve(m1, s.b, s.tol, first, oracle.run("cg", m1, s.b, prec=prec, niter=800, tol=s.tol))

    # the same pattern with new values: no rebuild, the oracle's answer on them
    for method in ("cg", "bicg", "cgs"):
        got = kr.solve_system(method, m2, b2, prec=prec, niter=800, tol=tol2)
        ref = oracle.run(method, m2, b2, prec=prec, niter=800, tol=tol2)
        print(f"prec={prec} {method} niter gpu={got['niter']} oracle={ref['niter']}", flush=True)
        _check_solve(m2, b2, tol2, got, ref)
    st2 = kr.dropin_stats()
    assert st2["pattern_builds"] == st1["pattern_builds"]
    assert st2["calls"] == st1["calls"] + 3

    # prec1_ and resid_ share the copy and stay bitwise the oracle's
    D = kr.prec1(m2.n, m2.nonz, m2.aval, m2.acol, m2.arow, m2.adia, prec)
    assert np.array_equal(D, oracle.prec1(m2.n, m2.nonz, m2.aval, m2.acol, m2.arow, m2.adia, prec))
    x = np.zeros(m2.n + 1)
    x[1:] = np.random.default_rng(74).standard_normal(m2.n)
    r_gpu, r_ref = b2.copy(), b2.copy()
    kr.resid(m2.n, 0, m2.nonz, m2.aval, m2.acol, m2.arow, np.zeros(1), x, r_gpu)
    oracle.resid(m2.n, 0, m2.nonz, m2.aval, m2.acol, m2.arow, np.zeros(1), x, r_ref)
    assert np.array_equal(r_gpu, r_ref)
    assert kr.dropin_stats()["pattern_builds"] == st1["pattern_builds"]

    # a solve on the kept copy gives the bits of a freshly built one
    kept = kr.solve_system("cg", m2, b2, prec=prec, niter=800, tol=tol2)
    kr.dropin_release()
    fresh = kr.solve_system("cg", m2, b2, prec=prec, niter=800, tol=tol2)
    assert kr.dropin_stats()["pattern_builds"] == st1["pattern_builds"] + 1
    assert kept["niter"] == fresh["niter"] and kept["fin_res"] == fresh["fin_res"]
    assert np.array_equal(kept["x"], fresh["x"]) and np.array_equal(kept["r"], fresh["r"])


def test_changed_pattern_is_rebuilt(pkg, oracle):
    kr = pkg.krylov
    mg = pkg.meshgen
    a = mg.pressure_system(10, seed=75)
    b = mg.pressure_system(10, seed=75, perm_seed=3)   # same N and NONZ, other Acol / Arow / Adia
    c = mg.pressure_system((12, 9, 8), seed=76)
    assert a.matrix.n == b.matrix.n and a.matrix.nonz == b.matrix.nonz
    assert not np.array_equal(a.matrix.arow, b.matrix.arow)
    kr.dropin_release()
    builds = kr.dropin_stats()["pattern_builds"]
    for k, s in enumerate((a, b, a, c, c)):
        got = kr.solve_system("cg", s.matrix, s.b, prec=1, niter=1000, tol=s.tol)
        ref = oracle.run("cg", s.matrix, s.b, prec=1, niter=1000, tol=s.tol)
        _check_solve(s.matrix, s.b, s.tol, got, ref)
        builds_now = kr.dropin_stats()["pattern_builds"]
        # a, b, a and c differ from the call before; the second c does not
        assert builds_now == builds + (1 if k < 4 else 0), k
        builds = builds_now
    # prec1_ on the pattern of the last solve keeps the copy
    m = c.matrix
    got = kr.prec1(m.n, m.nonz, m.aval, m.acol, m.arow, m.adia, 1)
    assert np.array_equal(got, oracle.prec1(m.n, m.nonz, m.aval, m.acol, m.arow, m.adia, 1))
    assert kr.dropin_stats()["pattern_builds"] == builds
    kr.dropin_release()
