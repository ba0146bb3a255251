"""FP32-inner / FP64-refinement solves against the FP64 oracle.

The mixed solver stops on the FP64 true residual ||b - A x|| < tol, so the
pass bar is that test evaluated independently here, plus the distance of x
to the oracle's FP64 solution, which both satisfy the same residual bound:
||x_mixed - x_oracle|| <= ||A^-1|| (||r_mixed|| + ||r_oracle||).  The relative
error against the oracle is printed per case.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inv_norm2(m):
    return 1.0 / np.linalg.svd(m.to_scipy().toarray(), compute_uv=False)[-1]


@pytest.mark.parametrize("case,method,prec", [
    ("pressure12", "cg", 1), ("pressure12", "cg", 2), ("pressure12", "cgs", 1),
    ("pressure10_perm", "cg", 1), ("momentum10", "cgs", 1), ("momentum10", "cgs", 2),
    ("pressure10_perm", "bicg", 1),
])
def test_mixed_matches_fp64_oracle(pkg, oracle, cuda, case, method, prec):
    import torch
    mg = pkg.meshgen
    s = {"pressure12": lambda: mg.pressure_system(12, seed=51),
         "pressure10_perm": lambda: mg.pressure_system(10, seed=52, perm_seed=53),
         "momentum10": lambda: mg.momentum_system(10, seed=54)}[case]()
    m = s.matrix
    A = m.to_scipy()
    ds = pkg.krylov.DeviceSystem(m)
    b = torch.from_numpy(s.b.copy()).to(cuda)
    x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
    r = torch.zeros_like(b)
    it, ini, fin, outer = ds.solve_mixed(method, prec, x, b, 5000, s.tol, inner_rtol=1e-3, r_out=r)
    xm = x.cpu().numpy()
    true_r = np.linalg.norm(s.b - A @ xm)
    assert np.array_equal(b.cpu().numpy(), s.b)           # b untouched
    assert ini == pytest.approx(np.linalg.norm(s.b), rel=1e-12)
    assert true_r < s.tol * (1 + 1e-9) and fin < s.tol
    assert fin == pytest.approx(true_r, rel=1e-6)
    assert np.allclose(r.cpu().numpy(), s.b - A @ xm, atol=1e-12 * np.linalg.norm(s.b))
    ref = oracle.run(method, m, s.b, prec=prec, niter=5000, tol=s.tol)
    assert ref["fin_res"] < s.tol
    rel = np.linalg.norm(xm - ref["x"]) / np.linalg.norm(ref["x"])
    rr = np.linalg.norm(s.b - A @ ref["x"])
    bound = _inv_norm2(m) * (true_r + rr) / np.linalg.norm(ref["x"])
    print(f"{case} {method} prec={prec}: inner={it} outer={outer} rel_err_vs_fp64={rel:.2e} bound={bound:.2e}")
    assert rel <= bound
    ds.close()


def test_mixed_already_converged_and_cap(pkg, cuda):
    import torch
    s = pkg.meshgen.pressure_system(8, seed=55)
    m = s.matrix
    ds = pkg.krylov.DeviceSystem(m)
    b = torch.from_numpy(s.b.copy()).to(cuda)
    # exact solution as the initial guess: no refinement step is taken
    xs = np.linalg.solve(m.to_scipy().toarray(), s.b)
    x = torch.from_numpy(xs.copy()).to(cuda)
    it, ini, fin, outer = ds.solve_mixed("cg", 1, x, b, 100, s.tol)
    assert outer == 0 and it == 0 and fin < s.tol
    # a cap that is too small returns count + 1 and a residual above tol
    x.zero_()
    it, ini, fin, outer = ds.solve_mixed("cg", 1, x, b, 3, s.tol)
    assert fin >= s.tol and it == 4
    ds.close()


def test_difference_form_fp32_operator(pkg, cuda):
    """The FP32 operator of the inner solves is A in difference form:
    (A p)_i = sum_{j!=i} a_ij (p_j - p_i) + rs_i p_i.  On a random vector it
    matches the FP64 product to FP32 rounding relative to |A||p|; on a
    constant vector it is exactly rs * c (rows of the pressure matrix sum
    to zero except at the pinned cell), which the plain FP32 product is not."""
    import torch
    s = pkg.meshgen.pressure_system(20, seed=56)
    m = s.matrix
    A = m.to_scipy()
    ds = pkg.krylov.DeviceSystem(m)
    p = np.random.default_rng(57).standard_normal(m.n).astype(np.float32)
    y = torch.zeros(m.n, dtype=torch.float32, device=cuda)
    ds.spmv32(torch.from_numpy(p).to(cuda), y)
    ref = A @ p.astype(np.float64)
    scale = abs(A) @ np.abs(p.astype(np.float64))
    assert np.all(np.abs(y.cpu().numpy() - ref) <= 16 * 6e-8 * scale + 1e-30)
    c = np.full(m.n, 0.75, np.float32)
    ds.spmv32(torch.from_numpy(c).to(cuda), y)
    rs = np.asarray(A.sum(axis=1)).ravel().astype(np.float32)
    got = y.cpu().numpy()
    diag = np.abs(A.diagonal())
    interior = np.abs(rs) < 1e-6 * diag
    # rows that sum to zero give zero to FP64 rounding of the row sum, far
    # below the eps32*|a_ii| that the plain FP32 product leaves
    assert np.all(np.abs(got[interior]) <= 1e-12 * diag[interior])
    assert np.allclose(got[~interior], rs[~interior] * np.float32(0.75), rtol=2.5e-7, atol=0)
    ds.close()


@pytest.mark.parametrize("case", ["pressure", "pressure_perm", "momentum"])
def test_sliced_fp32_operator_matches_csr_form(pkg, cuda, monkeypatch, case):
    """The FP32 inner operator in sliced ELL layout sums each row's terms in
    the same order as the CSR difference form (TFLOWS_FP32_OP=sliced / =csr,
    read when a system is created); its padding adds
    0 * (p_i - p_i).  Products, and the mixed solves built on them, agree
    with the CSR form's (== comparison: the padding can only turn a -0 row
    sum into +0)."""
    import torch
    mg = pkg.meshgen
    s = {"pressure": lambda: mg.pressure_system((23, 9, 7), seed=58),
         "pressure_perm": lambda: mg.pressure_system(11, seed=59, perm_seed=5),
         "momentum": lambda: mg.momentum_system(10, seed=60)}[case]()
    m = s.matrix
    p = torch.from_numpy(np.random.default_rng(61).standard_normal(m.n).astype(np.float32)).to(cuda)
    b = torch.from_numpy(s.b).to(cuda)
    method = "cgs" if case == "momentum" else "cg"
    out = {}
    for form in ("csr", "sliced"):
        monkeypatch.setenv("TFLOWS_FP32_OP", form)
        ds = pkg.krylov.DeviceSystem(m)
        y = torch.zeros(m.n, dtype=torch.float32, device=cuda)
        ds.spmv32(p, y)
        x1 = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        r1 = ds.solve_mixed(method, 1, x1, b, 2000, s.tol, inner_rtol=1e-3)
        x2 = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        r2 = ds.solve_mixed_reliable(1, x2, b, 2000, s.tol, interval=25) if method == "cg" else None
        torch.cuda.synchronize()
        out[form] = (y.cpu().numpy(), r1, x1.cpu().numpy(), r2, x2.cpu().numpy())
        ds.close()
    a, c = out["csr"], out["sliced"]
    assert np.array_equal(a[0], c[0])
    assert a[1] == c[1] and np.array_equal(a[2], c[2])
    assert a[3] == c[3] and np.array_equal(a[4], c[4])
    assert c[1][2] < s.tol


@pytest.mark.parametrize("case", ["pressure", "pressure_perm_wide"])
def test_fp32_sliced_column_width(pkg, cuda, monkeypatch, case):
    """The sliced FP32 operator holds its columns as 16-bit offsets from the
    row where every offset fits (TFLOWS_FP32_COL16=0 keeps 32-bit columns):
    the same columns, so the same products and mixed solves bit for bit.  A
    randomly renumbered system of 64,000 rows stored in the caller's order
    (TFLOWS_REORDER=0; the default locality order would bring its offsets
    within 16 bits) keeps 32-bit columns."""
    import torch
    mg = pkg.meshgen
    s = {"pressure": lambda: mg.pressure_system((23, 9, 7), seed=62),
         "pressure_perm_wide": lambda: mg.pressure_system(40, seed=63, perm_seed=6)}[case]()
    m = s.matrix
    p = torch.from_numpy(np.random.default_rng(64).standard_normal(m.n).astype(np.float32)).to(cuda)
    b = torch.from_numpy(s.b).to(cuda)
    out = {}
    monkeypatch.setenv("TFLOWS_FP32_OP", "sliced")
    monkeypatch.setenv("TFLOWS_SHARED_SLOTS", "0")       # per-entry columns, the layout under test
    if case == "pressure_perm_wide":
        monkeypatch.setenv("TFLOWS_REORDER", "0")
    for flag in ("0", "1"):
        monkeypatch.setenv("TFLOWS_FP32_COL16", flag)
        ds = pkg.krylov.DeviceSystem(m)
        y = torch.zeros(m.n, dtype=torch.float32, device=cuda)
        ds.spmv32(p, y)
        lay = ds.fp32_layout()
        x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        res = ds.solve_mixed_reliable(1, x, b, 3000, s.tol, interval=50)
        torch.cuda.synchronize()
        out[flag] = (lay, y.cpu().numpy(), res, x.cpu().numpy())
        ds.close()
    assert out["0"][0]["col_bytes"] == 4
    assert out["1"][0]["col_bytes"] == (2 if case == "pressure" else 4)
    assert np.array_equal(out["0"][1], out["1"][1])
    # iterations, residual norms and x bit for bit; the replacement count is
    # the host's (it may queue one more chunk before it sees the end)
    assert out["0"][2][:3] == out["1"][2][:3] and np.array_equal(out["0"][3], out["1"][3])
    assert out["1"][2][2] < s.tol


@pytest.mark.parametrize("case,prec,interval", [
    ("pressure12", 1, 50), ("pressure12", 0, 7), ("pressure10_perm", 1, 1), ("pressure10_perm", 1, 1000),
    ("momentum10", 1, 20),
    # DIC: the generic FP32 CG with FP32 sweeps, a replacement every
    # `interval` iterations
    ("pressure12", 2, 50), ("pressure10_perm", 2, 7), ("wall20", 2, 25),
])
def test_mixed_reliable_matches_fp64_oracle(pkg, oracle, cuda, case, prec, interval):
    """FP32 CG with FP64 residual replacement: the FP64 true residual meets
    tol, the iteration count stays near the FP64 CG's, and x is within the
    conditioning bound of the oracle's FP64 solution."""
    import torch
    mg = pkg.meshgen
    s = {"pressure12": lambda: mg.pressure_system(12, seed=51),
         "pressure10_perm": lambda: mg.pressure_system(10, seed=52, perm_seed=53),
         "momentum10": lambda: mg.momentum_system(10, seed=54),
         "wall20": lambda: mg.wall_pressure_system(20, seed=55)}[case]()
    m = s.matrix
    A = m.to_scipy()
    ds = pkg.krylov.DeviceSystem(m)
    b = torch.from_numpy(s.b.copy()).to(cuda)
    x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
    r = torch.zeros_like(b)
    cap = 3000
    it, ini, fin, rep = ds.solve_mixed_reliable(prec, x, b, cap, s.tol, interval=interval, r_out=r)
    xm = x.cpu().numpy()
    true_r = np.linalg.norm(s.b - A @ xm)
    assert ini == pytest.approx(np.linalg.norm(s.b), rel=1e-12)
    if case == "momentum10":
        # CG on the nonsymmetric momentum matrix stalls, as the reference's
        # does; the solver must report it (count + 1, residual above tol)
        ref = oracle.run("cg", m, s.b, prec=prec, niter=cap, tol=s.tol)
        assert ref["fin_res"] >= s.tol and fin >= s.tol and it == cap + 1
        assert fin == pytest.approx(true_r, rel=1e-6)
        ds.close()
        return
    assert true_r < s.tol and fin < s.tol and fin == pytest.approx(true_r, rel=1e-6)
    assert np.allclose(r.cpu().numpy(), s.b - A @ xm, atol=1e-12 * np.linalg.norm(s.b))
    ref = oracle.run("cg", m, s.b, prec=prec, niter=cap, tol=s.tol)
    rel = np.linalg.norm(xm - ref["x"]) / np.linalg.norm(ref["x"])
    bound = _inv_norm2(m) * (true_r + np.linalg.norm(s.b - A @ ref["x"])) / np.linalg.norm(ref["x"])
    print(f"{case} prec={prec} interval={interval}: it={it} (fp64 {ref['niter']}) replacements={rep} "
          f"rel_err_vs_fp64={rel:.2e} bound={bound:.2e}")
    assert rel <= bound
    assert it <= 1.5 * ref["niter"] + interval + 5
    ds.close()


@pytest.mark.parametrize("case", ["pressure12", "wall20", "momentum10"])
@pytest.mark.parametrize("interval", [1, 7, 50])
def test_mixed_reliable_carry_bitwise(pkg, cuda, case, interval):
    """The FP32 iterations carrying x += alfa p into the next p kernel, with
    the FP64 merge applying the update a pause or a replacement leaves
    pending, give the bits of the update kernel applying it
    (TFLOWS_CG_CARRY=0): converged solves and the capped momentum solve."""
    import os
    import torch
    mg = pkg.meshgen
    s = {"pressure12": lambda: mg.pressure_system(12, seed=51),
         "momentum10": lambda: mg.momentum_system(10, seed=54),
         "wall20": lambda: mg.wall_pressure_system(20, seed=55)}[case]()
    m = s.matrix
    ds = pkg.krylov.DeviceSystem(m)
    cap = 400 if case == "momentum10" else 3000
    out = {}
    for carry in ("0", "1"):
        os.environ["TFLOWS_CG_CARRY"] = carry
        try:
            b = torch.from_numpy(s.b.copy()).to(cuda)
            x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
            r = torch.zeros_like(b)
            res = ds.solve_mixed_reliable(1, x, b, cap, s.tol, interval=interval, r_out=r)
        finally:
            os.environ.pop("TFLOWS_CG_CARRY", None)
        out[carry] = (res, x.cpu().numpy(), r.cpu().numpy())
    a, c = out["0"], out["1"]
    print(f"case {case} interval {interval}: {a[0]} / {c[0]}", flush=True)
    # niter, IniRes, FinRes; the replacement count is the number of chunks
    # the host had issued when it saw the stop, which depends on timing
    assert a[0][:3] == c[0][:3]
    assert np.array_equal(a[1], c[1]) and np.array_equal(a[2], c[2])
    if case != "momentum10":
        assert c[0][2] < s.tol
    ds.close()
