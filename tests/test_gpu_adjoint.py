"""Adjoint solve + outer-product gradient against finite differences and
an FP64 CPU reference gradient (scipy direct solves)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_transpose_values_and_outer_grad(pkg, cuda):
    import torch
    s = pkg.meshgen.momentum_system(7, seed=61)          # nonsymmetric
    m = s.matrix
    ds = pkg.krylov.DeviceSystem(m)
    at = torch.zeros(m.nnz, dtype=torch.float64, device=cuda)
    ds.transpose_values(at)
    assert np.array_equal(at.cpu().numpy(), m.transpose_values()[: m.nnz])
    rng = np.random.default_rng(62)
    lam, x = rng.standard_normal(m.n), rng.standard_normal(m.n)
    g = torch.full((m.nonz,), 7.0, dtype=torch.float64, device=cuda)
    ds.outer_grad(torch.from_numpy(lam).to(cuda), torch.from_numpy(x).to(cuda), g)
    rows = m.row_of_entry()
    cols = m.arow[: m.nnz] - 1
    want = np.zeros(m.nonz)
    want[: m.nnz] = -lam[rows] * x[cols]
    assert np.array_equal(g.cpu().numpy(), want)
    ds.close()


@pytest.mark.parametrize("method,prec", [("cgs", 1), ("bicg", 1)])
def test_adjoint_solve_nonsymmetric(pkg, cuda, method, prec):
    """A^T lambda = g for a nonsymmetric matrix (CGS converges on it; BiCG
    with the reference's A-shadow stalls, reported as cap + 1)."""
    import torch
    s = pkg.meshgen.momentum_system(9, seed=63)
    m = s.matrix
    ds = pkg.krylov.DeviceSystem(m)
    g = np.random.default_rng(64).standard_normal(m.n)
    lam = torch.zeros(m.n, dtype=torch.float64, device=cuda)
    r = torch.from_numpy(g.copy()).to(cuda)
    tol = 1e-10 * np.linalg.norm(g)
    it, ini, fin = ds.solve_adjoint(method, prec, lam, r, 2000, tol)
    AT = m.to_scipy().T.tocsr()
    true_r = np.linalg.norm(g - AT @ lam.cpu().numpy())
    if method == "cgs":
        assert fin < tol and true_r < 10 * tol
    else:
        assert it == 2001 or fin < tol
    ds.close()


def test_gradients_match_fd_and_cpu_reference(pkg, cuda):
    from tflows_amd import adjoint_harness
    out = adjoint_harness.run(pkg, n=14, seed=4, n_check=50, h=1e-6, rtol=1e-13)
    print(out)
    assert out["ref_rel_grad_b"] < 1e-9
    assert out["ref_rel_grad_coef"] < 1e-9
    # 1e-6 relative against central differences (h = 1e-6) over the checked
    # entries as a vector; per entry within 1e-6 relative plus the
    # difference's own error bound (||lambda|| rtol ||b|| + eps L) / h
    assert out["fd_rel"] < 1e-6
    assert out["fd_floor_ratio"] <= 1.0
