"""GPU tests of the level-major DIC sweeps.

The DIC apply (Process/CG.f90:135-156) sweeps a level-major copy of the
triangles: rows of one forward level sit at consecutive positions and keep
their entries in matrix order, so every z_i is the same sum of the same
products as in the sweeps over the matrix rows.  Solves with and without
the copy (TFLOWS_DIC_PLAN=0, read when a system first builds its levels)
must agree bit for bit: single systems in natural and locality order, the
three methods, FP32 inner solves, batches, and after a value update.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(pkg, monkeypatch, make):
    """Two copies of an object whose levels are built with and without the
    level-major copy; `make(plan)` creates one and runs its first DIC use."""
    monkeypatch.setenv("TFLOWS_DIC_PLAN", "0")
    old = make(False)
    monkeypatch.delenv("TFLOWS_DIC_PLAN")
    new = make(True)
    return old, new


def _system(pkg, case):
    mg = pkg.meshgen
    if case == "natural":
        return mg.pressure_system(14, seed=61, prec=2)
    if case == "renumbered":
        return mg.pressure_system(12, seed=62, perm_seed=7, prec=2)
    return mg.momentum_system(11, seed=63, prec=2)


@pytest.mark.parametrize("case", ["natural", "renumbered", "momentum"])
@pytest.mark.parametrize("method", ["cg", "bicg", "cgs"])
def test_plan_solves_bitwise(pkg, cuda, monkeypatch, case, method):
    import torch
    s = _system(pkg, case)
    m = s.matrix
    runs = {}

    def make(plan):
        ds = pkg.krylov.DeviceSystem(m)
        out = []
        for cap in (7, 400):
            x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
            r = torch.from_numpy(s.b.copy()).to(cuda)
            it, ini, fin = ds.solve(method, 2, x, r, cap, s.tol)
            torch.cuda.synchronize()
            out.append((it, ini, fin, x.cpu().numpy(), r.cpu().numpy()))
        runs[plan] = out
        return ds

    old, new = _pair(pkg, monkeypatch, make)
    for a, b in zip(runs[False], runs[True]):
        assert a[:3] == b[:3]
        assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
    if case != "momentum" or method == "cgs":
        assert runs[True][1][2] < s.tol
    old.close()
    new.close()


@pytest.mark.parametrize("method", ["cg", "cgs"])
def test_plan_after_value_update(pkg, cuda, monkeypatch, method):
    """New values reach the level-major copy before the next sweep."""
    import torch
    s = _system(pkg, "momentum")
    m = s.matrix
    scale = torch.from_numpy(1.0 + 0.25 * np.random.default_rng(3).random(m.nonz)).to(cuda)
    runs = {}

    def make(plan):
        ds = pkg.krylov.DeviceSystem(m)
        out = []
        for step in range(2):
            if step:
                ds.set_values(ds.values * scale)
            x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
            r = torch.from_numpy(s.b.copy()).to(cuda)
            it, ini, fin = ds.solve(method, 2, x, r, 400, s.tol)
            torch.cuda.synchronize()
            out.append((it, fin, x.cpu().numpy()))
        runs[plan] = out
        return ds

    old, new = _pair(pkg, monkeypatch, make)
    for a, b in zip(runs[False], runs[True]):
        assert a[:2] == b[:2] and np.array_equal(a[2], b[2])
    assert not np.array_equal(runs[True][0][2], runs[True][1][2])
    old.close()
    new.close()


@pytest.mark.parametrize("method", ["cg", "cgs"])
def test_plan_mixed_bitwise(pkg, cuda, monkeypatch, method):
    """FP32 inner solves sweep the FP32 copy, rounded from the same FP64
    values as the FP32 matrix."""
    import torch
    s = _system(pkg, "momentum" if method == "cgs" else "renumbered")
    m = s.matrix
    runs = {}

    def make(plan):
        ds = pkg.krylov.DeviceSystem(m)
        x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        b = torch.from_numpy(s.b.copy()).to(cuda)
        res = ds.solve_mixed(method, 2, x, b, 2000, s.tol, inner_rtol=1e-3)
        torch.cuda.synchronize()
        runs[plan] = (res, x.cpu().numpy())
        return ds

    old, new = _pair(pkg, monkeypatch, make)
    assert runs[False][0] == runs[True][0]
    assert np.array_equal(runs[False][1], runs[True][1])
    assert runs[True][0][2] < s.tol
    old.close()
    new.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_plan_batch_bitwise(pkg, cuda, monkeypatch, mixed):
    """Batches sweep the whole batch's levels; fixed grid and lanes so that
    both runs use the same reduction order."""
    import torch
    systems = pkg.meshgen.transport_batch(n=10, count=4, seed=5, prec=2)
    runs = {}

    def make(plan):
        B = pkg.krylov.DeviceBatch([t.matrix for t in systems])
        B.configure(grid=1024, lanes=1, chunk=1, rebalance=True)
        bvec = np.concatenate([t.b for t in systems])
        tol = np.array([t.tol for t in systems])
        x = torch.zeros(B.rows, dtype=torch.float64, device=cuda)
        if mixed:
            b = torch.from_numpy(bvec).to(cuda)
            res = B.solve_mixed("cgs", 2, x, b, 2000, tol, inner_rtol=1e-3)
            r = b
        else:
            r = torch.from_numpy(bvec).to(cuda)
            res = B.solve("cgs", 2, x, r, 2000, tol)
        torch.cuda.synchronize()
        runs[plan] = (res, x.cpu().numpy(), r.cpu().numpy())
        return B

    old, new = _pair(pkg, monkeypatch, make)
    ra, rb = runs[False][0], runs[True][0]
    for key in ("niter", "fin_res"):
        assert np.array_equal(ra[key], rb[key])
    assert np.array_equal(runs[False][1], runs[True][1])
    assert np.array_equal(runs[False][2], runs[True][2])
    assert np.all(rb["fin_res"] < np.array([t.tol for t in systems]))
    old.close()
    new.close()
