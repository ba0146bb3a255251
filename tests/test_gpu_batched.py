"""GPU tests of the batched solver (many independent systems per launch)
and of the lane-group SpMV forms.

Every system of a batch must follow the reference loop on its own matrix:
the per-system results are checked against the oracle (bitwise-pinned to
the reference) to the same bar as single solves -- x within 1e-10 relative
L2 where the conditioning allows, the same stopping state, niter within
one -- and against single-system GPU solves.  The SpMV forms (1, 2, 4, 8
lanes per row) sum every row in column order, so they must agree bit for
bit with each other and with the oracle's SpMV.  The dot products fused
into the SpMV kernels are summed by whichever lane finishes a row, so a
whole solve changes with the form only by the rounding of those sums.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def mixed_sizes(pkg):
    """Six systems of uneven sizes and kinds, one smaller than a tile and
    one with a single row."""
    mg = pkg.meshgen
    out = [
        mg.pressure_system(12, seed=21),
        mg.pressure_system((9, 14, 5), seed=22, perm_seed=3),
        mg.pressure_system((20, 6, 3), seed=23, ratio=1.3),
        mg.pressure_system((5, 4, 3), seed=24),
        mg.pressure_system(16, seed=25),
    ]
    one = mg.from_scipy(__import__("scipy.sparse", fromlist=["csr_matrix"]).csr_matrix(np.array([[4.0]])))
    b1 = np.array([2.0])
    out.append(mg.LinearSystem(one, b1, np.zeros(1), 1e-8 * 2.0, 1, 100, {}))
    return out


def _batch_solve(pkg, cuda, systems, method, prec, cap, grid=0, lanes=0, chunk=0, rebalance=True,
                 fin_in=0.0, batch=None):
    import torch
    B = batch or pkg.krylov.DeviceBatch([s.matrix for s in systems])
    B.configure(grid=grid, lanes=lanes, chunk=chunk, rebalance=rebalance)
    b = np.concatenate([s.b for s in systems])
    x = torch.zeros(B.rows, dtype=torch.float64, device=cuda)
    r = torch.from_numpy(b).to(cuda)
    tol = np.array([s.tol for s in systems])
    res = B.solve(method, prec, x, r, cap, tol, fin_in=fin_in)
    torch.cuda.synchronize()
    xs = [v.cpu().numpy() for v in B.split(x)]
    rs = [v.cpu().numpy() for v in B.split(r)]
    return B, res, xs, rs


@pytest.mark.parametrize("method", ["cg", "bicg", "cgs"])
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_batch_matches_oracle(pkg, oracle, cuda, mixed_sizes, method, prec):
    systems = mixed_sizes
    cap = 1000
    B, res, xs, rs = _batch_solve(pkg, cuda, systems, method, prec, cap)
    for k, s in enumerate(systems):
        ref = oracle.run(method, s.matrix, s.b, prec=prec, niter=cap, tol=s.tol)
        print(f"system {k}: n={s.matrix.n} niter gpu={res['niter'][k]} ref={ref['niter']}", flush=True)
        assert abs(int(res["niter"][k]) - ref["niter"]) <= 1
        assert res["fin_res"][k] < s.tol
        assert ref["fin_res"] < s.tol
        assert res["ini_res"][k] == pytest.approx(ref["ini_res"], rel=1e-13)
        # the true residual of the returned x meets the tolerance with margin
        true_r = s.b - s.matrix.to_scipy() @ xs[k]
        assert np.linalg.norm(true_r) < 10 * s.tol
        # final residual vector returned in r1
        assert np.linalg.norm(rs[k]) == pytest.approx(res["fin_res"][k], rel=1e-10)
        if s.matrix.n > 1:
            A = s.matrix.to_scipy().toarray()
            inv = 1.0 / np.linalg.svd(A, compute_uv=False)[-1]
            bound = inv * (np.linalg.norm(rs[k]) + np.linalg.norm(ref["r"]))
            assert (_rel(xs[k], ref["x"]) <= 1e-10 or
                    np.linalg.norm(xs[k] - ref["x"]) <= 1.01 * bound)


def test_batch_trajectory_matches_oracle(pkg, oracle, cuda, mixed_sizes):
    """Capped runs: every system stops at cap + 1 with iterates equal to the
    oracle's to rounding."""
    cap = 5
    systems = mixed_sizes[:5]
    for method in ("cg", "bicg", "cgs"):
        B, res, xs, rs = _batch_solve(pkg, cuda, systems, method, 1, cap)
        for k, s in enumerate(systems):
            ref = oracle.run(method, s.matrix, s.b, prec=1, niter=cap, tol=s.tol)
            assert ref["niter"] == cap + 1
            assert int(res["niter"][k]) == cap + 1
            tol_x = 1e-9 if method == "cgs" else 1e-11
            assert _rel(xs[k], ref["x"]) <= tol_x


def test_batch_per_system_caps_and_early_exit(pkg, oracle, cuda, mixed_sizes):
    """Per-system caps; a zero right-hand side takes the ||b|| < tol exit
    (niter 0, IniRes unwritten, FinRes = the value passed in)."""
    import torch
    systems = mixed_sizes[:4]
    B = pkg.krylov.DeviceBatch([s.matrix for s in systems])
    b = [s.b.copy() for s in systems]
    b[2][:] = 0.0
    x = torch.zeros(B.rows, dtype=torch.float64, device=cuda)
    r = torch.from_numpy(np.concatenate(b)).to(cuda)
    caps = np.array([3, 1000, 50, 0], np.int32)
    tol = np.array([s.tol for s in systems])
    res = B.solve("cg", 1, x, r, caps, tol, fin_in=np.array([7.0, 7.0, 0.125, 7.0]))
    assert res["niter"][0] == 4                      # cap + 1
    assert res["niter"][2] == 0 and np.isnan(res["ini_res"][2]) and res["fin_res"][2] == 0.125
    assert res["niter"][3] == 1                      # cap <= 0: niter = 1 (CG.f90:259-262)
    ref = oracle.run("cg", systems[1].matrix, systems[1].b, prec=1, niter=1000, tol=systems[1].tol)
    assert abs(int(res["niter"][1]) - ref["niter"]) <= 1
    ref0 = oracle.run("cg", systems[0].matrix, systems[0].b, prec=1, niter=3, tol=systems[0].tol)
    x0 = B.split(x)[0].cpu().numpy()
    assert _rel(x0, ref0["x"]) <= 1e-11


def _fp64_op(monkeypatch, form):
    """Systems created from here on use the FP64 operator `form`: "csr"
    (the LDS-staged CSR kernels, which have the lane-group forms) or
    "sliced" (the sliced ELL kernels, one form)."""
    monkeypatch.setenv("TFLOWS_FP64_OP", form)


@pytest.mark.parametrize("form", ["csr", "sliced"])
def test_spmv_lane_forms_bitwise(pkg, oracle, cuda, monkeypatch, form):
    """q = A p is bitwise identical for 1, 2, 4, 8 lanes per row (CSR
    operator) and for the sliced operator, and equal to the oracle's
    sequential row sums (CG.f90:193-199); the residual form likewise."""
    import torch
    _fp64_op(monkeypatch, form)
    mg = pkg.meshgen
    for s in (mg.pressure_system(24, seed=31, perm_seed=5), mg.momentum_system(17, seed=32)):
        m = s.matrix
        p = np.random.default_rng(3).standard_normal(m.n)
        want = oracle.spmv(m, p)
        D = pkg.krylov.DeviceSystem(m)
        pt = torch.from_numpy(p).to(cuda)
        for lanes in (1, 2, 4, 8):
            D.set_lanes(lanes)
            q = torch.empty_like(pt)
            D.spmv(pt, q)
            r = torch.from_numpy(s.b.copy()).to(cuda)
            D.resid(pt, r)
            torch.cuda.synchronize()
            assert np.array_equal(q.cpu().numpy(), want), lanes
            rr = s.b.copy()
            xx = np.zeros(m.n + 1)
            xx[1:] = p
            oracle.resid(m.n, 0, m.nonz, m.aval, m.acol, m.arow, None, xx, rr)
            assert np.array_equal(r.cpu().numpy(), rr), lanes
        D.close()


@pytest.mark.parametrize("form", ["csr", "sliced"])
def test_single_solve_lane_forms(pkg, oracle, cuda, monkeypatch, form):
    """A single-system solve in every SpMV form: the same stopping state and
    x within the parity bar of the oracle; the same form twice gives the
    same bits."""
    import torch
    _fp64_op(monkeypatch, form)
    s = pkg.meshgen.pressure_system(20, seed=33)
    ref = oracle.run("cg", s.matrix, s.b, prec=1, niter=1000, tol=s.tol)
    D = pkg.krylov.DeviceSystem(s.matrix)
    repeats = []
    for lanes in (1, 2, 4, 8, 8):
        D.set_lanes(lanes)
        x = torch.zeros(s.matrix.n, dtype=torch.float64, device=cuda)
        r = torch.from_numpy(s.b.copy()).to(cuda)
        it, ini, fin = D.solve("cg", 1, x, r, 1000, s.tol)
        xv = x.cpu().numpy()
        assert abs(it - ref["niter"]) <= 1 and fin < s.tol
        assert _rel(xv, ref["x"]) <= 1e-10
        if lanes == 8:
            repeats.append(xv)
    assert np.array_equal(repeats[0], repeats[1])
    D.close()


@pytest.mark.parametrize("form", ["csr", "sliced"])
def test_batch_lane_forms(pkg, oracle, cuda, mixed_sizes, monkeypatch, form):
    """A batched solve in every SpMV form stays within the parity bar;
    repeated solves in one configuration (graph replays) give identical
    bits."""
    _fp64_op(monkeypatch, form)
    systems = mixed_sizes[:5]
    refs = [oracle.run("bicg", s.matrix, s.b, prec=1, niter=1000, tol=s.tol) for s in systems]
    B = None
    for lanes in (1, 2, 4, 8):
        B, res, xs, _ = _batch_solve(pkg, cuda, systems, "bicg", 1, 1000, grid=64, lanes=lanes,
                                     rebalance=False, batch=B)
        for k, s in enumerate(systems):
            assert abs(int(res["niter"][k]) - refs[k]["niter"]) <= 1
            assert res["fin_res"][k] < s.tol
            A = s.matrix.to_scipy().toarray()
            inv = 1.0 / np.linalg.svd(A, compute_uv=False)[-1]
            assert (_rel(xs[k], refs[k]["x"]) <= 1e-10 or
                    np.linalg.norm(xs[k] - refs[k]["x"]) <= 1.01 * inv * (res["fin_res"][k] + refs[k]["fin_res"]))
    first = np.concatenate(xs)
    for _ in range(2):
        B, res, xs, _ = _batch_solve(pkg, cuda, systems, "bicg", 1, 1000, grid=64, lanes=8, rebalance=False,
                                     batch=B)
        assert np.array_equal(np.concatenate(xs), first)


def test_batch_rebalance_reproducible(pkg, cuda, mixed_sizes):
    """Re-mapping the blocks of finished systems is decided on the state at
    a chunk boundary (a snapshot the last chunk writes), never on how far
    the device has run when the host looks: the same batch solved twice,
    rebalancing on the way, gives the same bits."""
    import scipy.sparse as sp
    mg = pkg.meshgen
    one = mg.from_scipy(sp.csr_matrix(np.array([[4.0]])))
    systems = list(mixed_sizes[:5]) + [mg.LinearSystem(one, np.array([2.0 + k]), np.zeros(1), 1e-8, 1, 100, {})
                                       for k in range(12)]
    runs = []
    for _ in range(3):
        B, res, xs, rs = _batch_solve(pkg, cuda, systems, "cgs", 1, 2000, grid=128, lanes=1, chunk=2,
                                      rebalance=True)
        runs.append((res, xs, B.config("cgs", 1)["rebalances"]))
        B.close()
    assert runs[0][2] >= 1
    for res, xs, reb in runs[1:]:
        assert reb == runs[0][2]
        assert np.array_equal(res["niter"], runs[0][0]["niter"])
        assert np.array_equal(res["fin_res"], runs[0][0]["fin_res"])
        for a, b in zip(xs, runs[0][1]):
            assert np.array_equal(a, b)


def test_batch_rebalance_and_chunks(pkg, oracle, cuda, mixed_sizes):
    """Handing blocks of finished systems to running ones, and different
    graph chunk sizes, leave every system's result within the parity bar.
    The batch is one 16^3 system plus twenty one-row systems that finish
    in one iteration, so the running share of the tiles drops below 70%
    and the blocks are re-mapped."""
    import scipy.sparse as sp
    mg = pkg.meshgen
    big = mixed_sizes[4]
    one = mg.from_scipy(sp.csr_matrix(np.array([[4.0]])))
    systems = [big] + [mg.LinearSystem(one, np.array([2.0 + k]), np.zeros(1), 1e-8, 1, 100, {})
                       for k in range(20)]
    ref = oracle.run("cg", big.matrix, big.b, prec=1, niter=1000, tol=big.tol)
    base = None
    for chunk, rebalance in ((8, False), (8, True), (1, True), (32, True)):
        B, res, xs, _ = _batch_solve(pkg, cuda, systems, "cg", 1, 1000, grid=64, lanes=0, chunk=chunk,
                                     rebalance=rebalance)
        cfg = B.config("cg", 1)
        print(f"chunk={chunk} rebalance={rebalance} cfg={cfg} niter={res['niter'].tolist()}", flush=True)
        if rebalance:
            assert cfg["rebalances"] >= 1
        else:
            assert cfg["rebalances"] == 0
        if base is None:
            base = (res["niter"].copy(), xs)
        assert abs(int(res["niter"][0]) - ref["niter"]) <= 1
        assert _rel(xs[0], ref["x"]) <= 1e-10
        for k, s in enumerate(systems):
            assert abs(int(res["niter"][k]) - int(base[0][k])) <= 1
            assert res["fin_res"][k] < s.tol
            assert _rel(xs[k], base[1][k]) <= 1e-10
        B.close()


def test_batch_matches_single_gpu_solves(pkg, cuda):
    """Transport systems on one shared pattern (the u/v/w/Zmix/Y batch at
    small size): CGS batched vs one GPU solve per system."""
    import torch
    systems = pkg.meshgen.transport_batch(n=14, count=5, seed=3)
    B, res, xs, _ = _batch_solve(pkg, cuda, systems, "cgs", 1, 500)
    for k, s in enumerate(systems):
        D = pkg.krylov.DeviceSystem(s.matrix)
        x = torch.zeros(s.matrix.n, dtype=torch.float64, device=cuda)
        r = torch.from_numpy(s.b.copy()).to(cuda)
        it, ini, fin = D.solve("cgs", 1, x, r, 500, s.tol)
        assert abs(int(res["niter"][k]) - it) <= 1
        assert res["fin_res"][k] < s.tol and fin < s.tol
        assert _rel(xs[k], x.cpu().numpy()) <= 1e-9
        D.close()


def _mixed(pkg, cuda, systems, method, prec, cap=4000):
    import torch
    B = pkg.krylov.DeviceBatch([s.matrix for s in systems])
    b = torch.from_numpy(np.concatenate([s.b for s in systems])).to(cuda)
    x = torch.zeros(B.rows, dtype=torch.float64, device=cuda)
    r_out = torch.zeros_like(x)
    tol = np.array([s.tol for s in systems])
    res = B.solve_mixed(method, prec, x, b, cap, tol, inner_rtol=1e-4, r_out=r_out)
    xs = [v.cpu().numpy() for v in B.split(x)]
    rs = [v.cpu().numpy() for v in B.split(r_out)]
    B.close()
    return res, xs, rs


@pytest.mark.parametrize("method,prec", [("cg", 1), ("cg", 2), ("bicg", 1), ("cgs", 1), ("cgs", 2)])
def test_batch_mixed_precision(pkg, oracle, cuda, mixed_sizes, method, prec):
    """FP32 inner solves with FP64 refinement reach each system's absolute
    tolerance on the FP64 true residual; x agrees with the FP64 oracle
    solution to the accuracy the tolerance fixes.  CG / BiCG on the pressure
    systems; CGS on the transport systems it is meant for (on the pinned
    pressure matrices FP32 CGS can diverge, see the next test)."""
    if method == "cgs":
        systems = pkg.meshgen.transport_batch(n=12, count=5, seed=3)
    else:
        systems = mixed_sizes
    res, xs, rs = _mixed(pkg, cuda, systems, method, prec)
    for k, s in enumerate(systems):
        assert res["fin_res"][k] < s.tol, k
        true_r = s.b - s.matrix.to_scipy() @ xs[k]
        assert np.linalg.norm(true_r) == pytest.approx(res["fin_res"][k], rel=1e-6, abs=1e-3 * s.tol)
        assert np.linalg.norm(rs[k]) == pytest.approx(res["fin_res"][k], rel=1e-12)
        ref = oracle.run(method, s.matrix, s.b, prec=prec, niter=4000, tol=s.tol)
        if s.matrix.n > 1:
            A = s.matrix.to_scipy().toarray()
            inv = 1.0 / np.linalg.svd(A, compute_uv=False)[-1]
            assert np.linalg.norm(xs[k] - ref["x"]) <= 1.01 * inv * (res["fin_res"][k] + ref["fin_res"])


def test_batch_mixed_guard(pkg, cuda, mixed_sizes):
    """FP32 CGS on the pinned pressure matrices can diverge (its squared
    residual polynomial peaks beyond what FP32 resolves).  The inner solve
    then stops on its divergence guard, the refinement step is not applied,
    and the system reports non-convergence the reference way: a finite x
    no worse than before the step, niter = iterations + 1, FinRes =
    ||b - A x|| of that x."""
    res, xs, rs = _mixed(pkg, cuda, mixed_sizes, "cgs", 1)
    for k, s in enumerate(mixed_sizes):
        assert np.all(np.isfinite(xs[k]))
        assert np.isfinite(res["fin_res"][k])
        assert res["fin_res"][k] <= res["ini_res"][k]
        true_r = s.b - s.matrix.to_scipy() @ xs[k]
        assert np.linalg.norm(true_r) == pytest.approx(res["fin_res"][k], rel=1e-6, abs=1e-3 * s.tol)
        print(f"system {k}: niter={res['niter'][k]} fin/tol={res['fin_res'][k] / s.tol:.3g}", flush=True)


# ---------------------------------------------------------------------------
# The persistent form: one launch per chunk of iterations, the kernels of an
# iteration as phases separated by barriers over each system's blocks.  At
# the same grid and lanes it sums the same tiles into the same reduction
# sites as the per-kernel form, so every result is the same bits.
# ---------------------------------------------------------------------------
def _form_solve(pkg, cuda, systems, form, method, prec, cap, grid, lanes=1, rebalance=True, chunk=0):
    B = pkg.krylov.DeviceBatch([s.matrix for s in systems])
    B.set_form(form)
    B, res, xs, rs = _batch_solve(pkg, cuda, systems, method, prec, cap, grid=grid, lanes=lanes, chunk=chunk,
                                  rebalance=rebalance, batch=B)
    used = B.config(method, prec)["form"]
    B.close()
    return used, res, xs, rs


def _same_results(a, b):
    for key in ("niter", "fin_res"):
        assert np.array_equal(a[1][key], b[1][key]), key
    assert np.array_equal(a[1]["ini_res"], b[1]["ini_res"], equal_nan=True)
    for u, v in zip(a[2] + a[3], b[2] + b[3]):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("method", ["cg", "bicg"])
@pytest.mark.parametrize("prec", [0, 1])
def test_persistent_form_bitwise(pkg, oracle, cuda, mixed_sizes, method, prec):
    systems = mixed_sizes
    per_kernel = _form_solve(pkg, cuda, systems, 1, method, prec, 1000, grid=256)
    persistent = _form_solve(pkg, cuda, systems, 2, method, prec, 1000, grid=256)
    assert per_kernel[0] == 1 and persistent[0] == 2
    _same_results(per_kernel, persistent)
    res, xs = persistent[1], persistent[2]
    for k, s in enumerate(systems):
        ref = oracle.run(method, s.matrix, s.b, prec=prec, niter=1000, tol=s.tol)
        assert abs(int(res["niter"][k]) - ref["niter"]) <= 1
        assert res["fin_res"][k] < s.tol
        assert np.linalg.norm(s.b - s.matrix.to_scipy() @ xs[k]) < 10 * s.tol


def test_persistent_form_caps_chunks_lanes(pkg, cuda, mixed_sizes):
    """Capped systems stop at cap + 1 inside a chunk; chunk lengths that do
    not divide the iteration counts, a lane-group SpMV form and a grid
    without rebalancing give the per-kernel form's bits."""
    systems = mixed_sizes[:5]
    for cap, chunk, lanes, grid, reb in ((7, 3, 1, 128, True), (1000, 5, 4, 256, True), (1000, 16, 2, 64, False)):
        a = _form_solve(pkg, cuda, systems, 1, "cg", 1, cap, grid, lanes=lanes, rebalance=reb, chunk=chunk)
        b = _form_solve(pkg, cuda, systems, 2, "cg", 1, cap, grid, lanes=lanes, rebalance=reb, chunk=chunk)
        assert b[0] == 2
        _same_results(a, b)
        if cap == 7:
            assert int(np.max(b[1]["niter"])) == cap + 1


def test_persistent_form_mixed_bitwise(pkg, cuda, mixed_sizes):
    """FP32 inner CG solves of the batched refinement in the persistent form."""
    import torch
    systems = mixed_sizes
    out = []
    for form in (1, 2):
        B = pkg.krylov.DeviceBatch([s.matrix for s in systems])
        B.set_form(form)
        B.configure(grid=256, lanes=1)
        b = torch.from_numpy(np.concatenate([s.b for s in systems])).to(cuda)
        x = torch.zeros(B.rows, dtype=torch.float64, device=cuda)
        res = B.solve_mixed("cg", 1, x, b, 2000, np.array([s.tol for s in systems]), inner_rtol=1e-3)
        torch.cuda.synchronize()
        assert B.config("cg", 1, bits=32)["form"] == form
        out.append((res, x.cpu().numpy()))
        B.close()
    assert np.array_equal(out[0][0]["niter"], out[1][0]["niter"])
    assert np.array_equal(out[0][0]["fin_res"], out[1][0]["fin_res"])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.all(out[1][0]["fin_res"] < np.array([s.tol for s in systems]))


def test_persistent_form_uneven_batch(pkg, cuda):
    """Twenty uneven pressure boxes (the many-system case) with rebalancing:
    the persistent and per-kernel forms agree bit for bit, and the tuned
    batch converges every system."""
    systems = pkg.meshgen.uneven_pressure_batch(count=20, seed=9, n_min=500, n_max=20000)
    a = _form_solve(pkg, cuda, systems, 1, "cg", 1, 5000, 512)
    b = _form_solve(pkg, cuda, systems, 2, "cg", 1, 5000, 512)
    assert b[0] == 2
    _same_results(a, b)
    used, res, xs, rs = _form_solve(pkg, cuda, systems, 0, "cg", 1, 5000, 0, lanes=0)
    print(f"tuned form {used}, iterations {int(np.sum(res['niter']))}", flush=True)
    assert np.all(res["fin_res"] < np.array([s.tol for s in systems]))


def test_persistent_wait_bound(pkg, cuda, mixed_sizes, monkeypatch):
    """Failure handling of the persistent form, forced from the host: with
    TFLOWS_BARRIER_WAIT_MS=0 (read when a batch is created) a block that
    has to wait at a barrier gives up at its first time check and raises
    the fault word, every other waiting block sees it and stops, so every
    wavefront ends; the batch then solves again in the per-kernel form from
    its inputs, keeps that form, and returns the per-kernel form's bits."""
    systems = mixed_sizes
    want = _form_solve(pkg, cuda, systems, 1, "cg", 1, 1000, 256)
    monkeypatch.setenv("TFLOWS_BARRIER_WAIT_MS", "0")
    got = _form_solve(pkg, cuda, systems, 2, "cg", 1, 1000, 256)
    monkeypatch.delenv("TFLOWS_BARRIER_WAIT_MS")
    print(f"wait bound 0: form after the solve = {got[0]}", flush=True)
    assert got[0] in (1, 2)
    _same_results(want, got)


def _grouped(pkg, monkeypatch, systems, budget_mb="0.01"):
    """A batch whose systems run in MALL-sized groups, with the group budget
    shrunk so that small test systems form several groups."""
    monkeypatch.setenv("TFLOWS_BATCH_GROUP_MB", budget_mb)
    B = pkg.krylov.DeviceBatch([s.matrix for s in systems])
    B.set_form(1)
    B.set_groups(2)
    return B


@pytest.mark.parametrize("method", ["cg", "bicg", "cgs"])
@pytest.mark.parametrize("prec", [0, 1])
def test_batch_groups_match_oracle(pkg, oracle, cuda, mixed_sizes, monkeypatch, method, prec):
    """Systems solved group after group (each group on the whole grid, its
    iterations counted from the group's start) follow the reference loop
    each: niter within one, the same stopping state, x within the parity
    bar."""
    systems = mixed_sizes
    cap = 1000
    B = _grouped(pkg, monkeypatch, systems)
    B, res, xs, rs = _batch_solve(pkg, cuda, systems, method, prec, cap, batch=B)
    groups = B.config(method, prec)["groups"]
    print(f"{method} prec {prec}: {groups} groups, niter {list(res['niter'])}", flush=True)
    assert groups >= 3
    for k, s in enumerate(systems):
        ref = oracle.run(method, s.matrix, s.b, prec=prec, niter=cap, tol=s.tol)
        assert abs(int(res["niter"][k]) - ref["niter"]) <= 1
        assert res["fin_res"][k] < s.tol
        assert res["ini_res"][k] == pytest.approx(ref["ini_res"], rel=1e-13)
        true_r = s.b - s.matrix.to_scipy() @ xs[k]
        assert np.linalg.norm(true_r) < 10 * s.tol
        if s.matrix.n > 1:
            A = s.matrix.to_scipy().toarray()
            inv = 1.0 / np.linalg.svd(A, compute_uv=False)[-1]
            bound = inv * (np.linalg.norm(rs[k]) + np.linalg.norm(ref["r"]))
            assert (_rel(xs[k], ref["x"]) <= 1e-10 or
                    np.linalg.norm(xs[k] - ref["x"]) <= 1.01 * bound)
    B.close()


def test_batch_groups_caps_early_exit_and_reproducible(pkg, oracle, cuda, mixed_sizes, monkeypatch):
    """Groups with per-system caps (a capped system stops at cap + 1 with
    the oracle's iterate) and a zero right-hand side in a later group (the
    ||b|| < tol exit; its group is skipped); the same batch solved twice
    gives the same bits."""
    import torch
    systems = mixed_sizes[:5]
    runs = []
    for _ in range(2):
        B = _grouped(pkg, monkeypatch, systems)
        b = [s.b.copy() for s in systems]
        b[3][:] = 0.0
        x = torch.zeros(B.rows, dtype=torch.float64, device=cuda)
        r = torch.from_numpy(np.concatenate(b)).to(cuda)
        caps = np.array([3, 1000, 1000, 50, 7], np.int32)
        tol = np.array([s.tol for s in systems])
        res = B.solve("cgs", 1, x, r, caps, tol, fin_in=np.array([7.0, 7.0, 7.0, 0.125, 7.0]))
        torch.cuda.synchronize()
        assert B.config("cgs", 1)["groups"] >= 3
        xs = [v.cpu().numpy() for v in B.split(x)]
        runs.append((res, xs))
        B.close()
    res, xs = runs[0]
    assert res["niter"][0] == 4 and res["niter"][4] == 8          # cap + 1
    assert res["niter"][3] == 0 and np.isnan(res["ini_res"][3]) and res["fin_res"][3] == 0.125
    for k in (0, 4):
        ref = oracle.run("cgs", systems[k].matrix, systems[k].b, prec=1, niter=int(caps[k]), tol=systems[k].tol)
        assert ref["niter"] == caps[k] + 1
        assert _rel(xs[k], ref["x"]) <= 1e-9
    for k in (1, 2):
        ref = oracle.run("cgs", systems[k].matrix, systems[k].b, prec=1, niter=1000, tol=systems[k].tol)
        assert abs(int(res["niter"][k]) - ref["niter"]) <= 1
    res2, xs2 = runs[1]
    assert np.array_equal(res["niter"], res2["niter"]) and np.array_equal(res["fin_res"], res2["fin_res"])
    for a, b2 in zip(xs, xs2):
        assert np.array_equal(a, b2)
