"""GPU tests of the locality reordering of a system's internal storage.

A renumbered mesh is stored in reverse Cuthill-McKee order with every row's
entries kept in their original order, so the SpMV, Resid and Prec1 (DIC
pivots included) stay bitwise equal to the oracle, solves keep the parity
bar, and vectors are exchanged in the caller's numbering.  TFLOWS_REORDER
(read when a system is created) forces the order either way here.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture
def forced(request):
    """Create systems with TFLOWS_REORDER set to the given value."""
    old = os.environ.get("TFLOWS_REORDER")

    def make(pkg, m, mode):
        os.environ["TFLOWS_REORDER"] = str(mode)
        try:
            return pkg.krylov.DeviceSystem(m)
        finally:
            if old is None:
                os.environ.pop("TFLOWS_REORDER", None)
            else:
                os.environ["TFLOWS_REORDER"] = old
    return make


@pytest.mark.parametrize("case,mode", [("natural", 1), ("renumbered", 1), ("renumbered", 2), ("renumbered", 0)])
def test_reordered_products_bitwise(pkg, oracle, cuda, forced, case, mode):
    import torch
    mg = pkg.meshgen
    s = mg.pressure_system(14, seed=41) if case == "natural" else mg.momentum_system(13, seed=42)
    m = s.matrix
    ds = forced(pkg, m, mode)
    assert ds.reordered == (mode == 1 or (mode == 2 and case == "renumbered"))
    p = np.random.default_rng(5).standard_normal(m.n)
    pt = torch.from_numpy(p).to(cuda)
    q = torch.empty_like(pt)
    ds.spmv(pt, q)
    assert np.array_equal(q.cpu().numpy(), oracle.spmv(m, p))
    r = torch.from_numpy(s.b.copy()).to(cuda)
    ds.resid(pt, r)
    rr = s.b.copy()
    xx = np.zeros(m.n + 1)
    xx[1:] = p
    oracle.resid(m.n, 0, m.nonz, m.aval, m.acol, m.arow, None, xx, rr)
    assert np.array_equal(r.cpu().numpy(), rr)
    for prec in (1, 2):
        d = torch.empty(m.n, dtype=torch.float64, device=cuda)
        ds.prec1(prec, d)
        torch.cuda.synchronize()
        assert np.array_equal(d.cpu().numpy(), oracle.prec1(m.n, m.nonz, m.aval, m.acol, m.arow, m.adia, prec))
    ds.close()


@pytest.mark.parametrize("method", ["cg", "cgs"])
@pytest.mark.parametrize("prec", [1, 2])
def test_reordered_solves_match_oracle(pkg, oracle, cuda, forced, method, prec):
    import torch
    s = pkg.meshgen.pressure_system(12, seed=43, perm_seed=9)
    m = s.matrix
    ds = forced(pkg, m, 1)
    assert ds.reordered
    ref = oracle.run(method, m, s.b, prec=prec, niter=1000, tol=s.tol)
    x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
    r = torch.from_numpy(s.b.copy()).to(cuda)
    it, ini, fin = ds.solve(method, prec, x, r, 1000, s.tol)
    assert abs(it - ref["niter"]) <= 1 and fin < s.tol
    assert ini == pytest.approx(ref["ini_res"], rel=1e-13)
    xv = x.cpu().numpy()
    A = m.to_scipy().toarray()
    inv = 1.0 / np.linalg.svd(A, compute_uv=False)[-1]
    assert (_rel(xv, ref["x"]) <= 1e-10 or
            np.linalg.norm(xv - ref["x"]) <= 1.01 * inv * (np.linalg.norm(r.cpu().numpy()) + np.linalg.norm(ref["r"])))
    # the residual returned is in the caller's order: after 3 iterations it
    # equals b - A x of the returned x up to recurrence rounding
    x3 = torch.zeros_like(x)
    r3 = torch.from_numpy(s.b.copy()).to(cuda)
    ds.solve(method, prec, x3, r3, 3, s.tol)
    true3 = s.b - m.to_scipy() @ x3.cpu().numpy()
    assert _rel(r3.cpu().numpy(), true3) <= 1e-10
    # mixed precision through the same boundary
    xm = torch.zeros_like(x)
    b = torch.from_numpy(s.b.copy()).to(cuda)
    rout = torch.zeros_like(x)
    itm, inim, finm, outer = ds.solve_mixed(method, prec, xm, b, 4000, s.tol, r_out=rout)
    assert finm < s.tol
    assert np.linalg.norm(rout.cpu().numpy() - (s.b - m.to_scipy() @ xm.cpu().numpy())) <= 1e-3 * s.tol
    ds.close()


def test_reordered_adjoint_pieces(pkg, cuda, forced):
    """Transposed values and the outer-product gradient are in the caller's
    entry order, and the adjoint solve solves A^T lambda = g."""
    import torch
    s = pkg.meshgen.momentum_system(11, seed=44)
    m = s.matrix
    ds = forced(pkg, m, 1)
    assert ds.reordered
    at = torch.empty(m.nnz, dtype=torch.float64, device=cuda)
    ds.transpose_values(at)
    assert np.array_equal(at.cpu().numpy(), m.transpose_values()[:m.nnz])
    rng = np.random.default_rng(7)
    lam, xv = rng.standard_normal(m.n), rng.standard_normal(m.n)
    g = torch.empty(m.nnz + 1, dtype=torch.float64, device=cuda)
    ds.outer_grad(torch.from_numpy(lam).to(cuda), torch.from_numpy(xv).to(cuda), g)
    rows = m.row_of_entry()
    cols = m.arow[:m.nnz] - 1
    assert np.array_equal(g.cpu().numpy()[:m.nnz], -(lam[rows] * xv[cols]))
    rhs = rng.standard_normal(m.n)
    lt = torch.zeros(m.n, dtype=torch.float64, device=cuda)
    r1 = torch.from_numpy(rhs.copy()).to(cuda)
    tol = 1e-10 * np.linalg.norm(rhs)
    it, ini, fin = ds.solve_adjoint("cgs", 1, lt, r1, 2000, tol)
    assert fin < tol
    At = m.to_scipy().T.tocsr()
    assert np.linalg.norm(rhs - At @ lt.cpu().numpy()) <= 10 * tol
    ds.close()
