"""GPU tests of the sliced ELL operators.

The FP64 iterations run their SpMVs on A in sliced ELL layout (64-row
slices, slot-major, each row's entries in CSR order, padding skipped); every
row sum is the reference loop's (Process/CG.f90:193-199) bit for bit, so
solves on the sliced operator and on the CSR arrays (TFLOWS_FP64_OP=csr /
=sliced, read when a system is created) agree bit for bit: every method and
preconditioner, natural and locality order, batches.  Unset, a system times
both forms on its first use and keeps the faster; its results are the same
bits whichever it keeps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(monkeypatch, make, with_auto=False):
    out = {}
    for form in ("csr", "sliced") + (("auto",) if with_auto else ()):
        if form == "auto":
            monkeypatch.delenv("TFLOWS_FP64_OP", raising=False)
        else:
            monkeypatch.setenv("TFLOWS_FP64_OP", form)
        out[form] = make()
    if with_auto:
        return out["csr"], out["sliced"], out["auto"]
    return out["csr"], out["sliced"]


@pytest.mark.parametrize("case", ["pressure", "pressure_perm", "momentum"])
@pytest.mark.parametrize("method", ["cg", "bicg", "cgs"])
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_sliced_solves_bitwise(pkg, oracle, cuda, monkeypatch, case, method, prec):
    import torch
    mg = pkg.meshgen
    s = {"pressure": lambda: mg.pressure_system((13, 7, 5), seed=71),
         "pressure_perm": lambda: mg.pressure_system(9, seed=72, perm_seed=4),
         "momentum": lambda: mg.momentum_system(8, seed=73)}[case]()
    m = s.matrix
    p = torch.from_numpy(np.random.default_rng(74).standard_normal(m.n)).to(cuda)

    def make():
        ds = pkg.krylov.DeviceSystem(m)
        q = torch.empty_like(p)
        ds.spmv(p, q)
        x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        r = torch.from_numpy(s.b.copy()).to(cuda)
        res = ds.solve(method, prec, x, r, 300, s.tol)
        torch.cuda.synchronize()
        out = (q.cpu().numpy(), res, x.cpu().numpy(), r.cpu().numpy())
        ds.close()
        return out

    a, b, c = _both(monkeypatch, make, with_auto=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(b[0], oracle.spmv(m, p.cpu().numpy()))
    assert a[1] == b[1] == c[1]
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    assert np.array_equal(a[2], c[2]) and np.array_equal(a[0], c[0])


def test_sliced_batch_bitwise(pkg, cuda, monkeypatch):
    """A batch on the sliced operator, with fixed grid and lanes, gives the
    CSR operator's bits."""
    import torch
    mg = pkg.meshgen
    systems = [mg.pressure_system(10, seed=75), mg.pressure_system((7, 11, 4), seed=76, perm_seed=2),
               mg.momentum_system(7, seed=77), mg.pressure_system((3, 2, 2), seed=78)]

    def make():
        B = pkg.krylov.DeviceBatch([t.matrix for t in systems])
        B.configure(grid=256, lanes=1, chunk=4, rebalance=True)
        x = torch.zeros(B.rows, dtype=torch.float64, device=cuda)
        r = torch.from_numpy(np.concatenate([t.b for t in systems])).to(cuda)
        res = B.solve("cgs", 1, x, r, 500, np.array([t.tol for t in systems]))
        torch.cuda.synchronize()
        out = (res, x.cpu().numpy(), r.cpu().numpy())
        B.close()
        return out

    a, b = _both(monkeypatch, make)
    for key in ("niter", "fin_res"):
        assert np.array_equal(a[0][key], b[0][key]), key
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert np.all(b[0]["fin_res"] < np.array([t.tol for t in systems]))


@pytest.mark.parametrize("case", ["pressure", "pressure_perm_wide", "momentum"])
def test_fp64_sliced_column_width(pkg, cuda, monkeypatch, case):
    """The sliced FP64 operator holds its columns as 16-bit offsets from the
    row where every offset fits, padding marked apart (TFLOWS_FP64_COL16=0
    keeps 32-bit columns): the same columns in the same order, so SpMV and
    solves agree bit for bit.  The randomly renumbered 64,000-row system in
    the caller's order (TFLOWS_REORDER=0) does not fit and keeps 32-bit
    columns under both settings."""
    import torch
    mg = pkg.meshgen
    s = {"pressure": lambda: mg.pressure_system((21, 8, 7), seed=81),
         "pressure_perm_wide": lambda: mg.pressure_system(40, seed=82, perm_seed=7),
         "momentum": lambda: mg.momentum_system(9, seed=83)}[case]()
    m = s.matrix
    p = torch.from_numpy(np.random.default_rng(84).standard_normal(m.n)).to(cuda)
    monkeypatch.setenv("TFLOWS_FP64_OP", "sliced")
    monkeypatch.setenv("TFLOWS_SHARED_SLOTS", "0")       # per-entry columns, the layout under test
    if case == "pressure_perm_wide":
        monkeypatch.setenv("TFLOWS_REORDER", "0")
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("TFLOWS_FP64_COL16", flag)
        ds = pkg.krylov.DeviceSystem(m)
        q = torch.empty_like(p)
        ds.spmv(p, q)
        x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        r = torch.from_numpy(s.b.copy()).to(cuda)
        res = ds.solve("cgs" if case == "momentum" else "cg", 1, x, r, 3000, s.tol)
        torch.cuda.synchronize()
        out[flag] = (q.cpu().numpy(), res, x.cpu().numpy(), r.cpu().numpy())
        ds.close()
    a, b = out["0"], out["1"]
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    assert b[1][2] < s.tol


@pytest.mark.parametrize("case,prec,interval", [("pressure", 1, 50), ("pressure_perm", 0, 9), ("pressure", 1, 1)])
def test_mixed_replacement_residual_bitwise(pkg, cuda, monkeypatch, case, prec, interval):
    """The FP64 residual of every replacement of the reliable-update solve
    runs on the FP64 operator of the system (r = b - A x: the products
    subtracted from b in CSR order in either layout), so the mixed solve
    gives the same iterations, norms, x and final residual whichever layout
    the system keeps."""
    import torch
    mg = pkg.meshgen
    s = {"pressure": lambda: mg.pressure_system((23, 9, 7), seed=81),
         "pressure_perm": lambda: mg.pressure_system(11, seed=82, perm_seed=83)}[case]()
    m = s.matrix
    b = torch.from_numpy(s.b.copy()).to(cuda)

    def run():
        ds = pkg.krylov.DeviceSystem(m)
        x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        r = torch.zeros_like(b)
        res = ds.solve_mixed_reliable(prec, x, b, 3000, s.tol, interval=interval, r_out=r)
        torch.cuda.synchronize()
        ds.close()
        return res, x.cpu().numpy(), r.cpu().numpy()

    a, c = _both(monkeypatch, run)
    assert a[0][:3] == c[0][:3]
    assert np.array_equal(a[1], c[1]) and np.array_equal(a[2], c[2])
    assert c[0][2] < s.tol


# layouts of the sliced operators: per-entry int32 columns, 16-bit offsets,
# and column offsets shared by each slice (one list per slice, a mask per row)
# and, for symmetric values, only the diagonal and upper slots stored with
# the lower ones read from their mirrors (TFLOWS_SYM_SLOTS)
# (the shared-offset loops over every slot at once, or two slots at a time:
# TFLOWS_SPMV_UNROLL)
LAYOUTS = {"int32": {"TFLOWS_SHARED_SLOTS": "0", "TFLOWS_FP64_COL16": "0", "TFLOWS_FP32_COL16": "0"},
           "col16": {"TFLOWS_SHARED_SLOTS": "0", "TFLOWS_FP64_COL16": "1", "TFLOWS_FP32_COL16": "1"},
           "shared_full": {"TFLOWS_SHARED_SLOTS": "1", "TFLOWS_SYM_SLOTS": "0", "TFLOWS_SYM_SLOTS32": "0",
                           "TFLOWS_FP64_COL16": "1", "TFLOWS_FP32_COL16": "1", "TFLOWS_SPMV_UNROLL": "1"},
           "shared_full_unroll2": {"TFLOWS_SHARED_SLOTS": "1", "TFLOWS_SYM_SLOTS": "0", "TFLOWS_SYM_SLOTS32": "0",
                                   "TFLOWS_FP64_COL16": "1", "TFLOWS_FP32_COL16": "1", "TFLOWS_SPMV_UNROLL": "0"},
           "shared_unroll2": {"TFLOWS_SHARED_SLOTS": "1", "TFLOWS_SYM_SLOTS": "1", "TFLOWS_SYM_SLOTS32": "1",
                              "TFLOWS_FP64_COL16": "1", "TFLOWS_FP32_COL16": "1", "TFLOWS_SPMV_UNROLL": "0"},
           "shared": {"TFLOWS_SHARED_SLOTS": "1", "TFLOWS_SYM_SLOTS": "1", "TFLOWS_SYM_SLOTS32": "1",
                      "TFLOWS_FP64_COL16": "1", "TFLOWS_FP32_COL16": "1", "TFLOWS_SPMV_UNROLL": "1"}}


@pytest.mark.parametrize("case", ["pressure", "channel_aniso", "pressure_perm"])
def test_shared_slot_layout_bitwise(pkg, oracle, cuda, monkeypatch, case):
    """The shared-offset layout holds each slot's column as one offset per
    slice and marks the slots each row has; a row's slots run in its CSR
    order, so the FP64 and FP32 SpMVs, the solves and the reliable-update
    mixed solve give the other layouts' bits (and the FP64 SpMV the
    oracle's).  It is chosen on naturally numbered structured meshes; the
    randomly renumbered system (locality order, columns no longer ascending
    within a row) keeps per-entry columns.  With symmetric values the shared
    layout stores only the diagonal and upper slots and reads each lower
    value from its mirror, in the same slot order: the same bits again."""
    import torch
    mg = pkg.meshgen
    # cubes: flat boxes are stored in their reverse Cuthill-McKee order,
    # which the shared layout cannot follow
    s = {"pressure": lambda: mg.pressure_system(20, seed=91),
         "channel_aniso": lambda: mg.wall_pressure_system(24),
         "pressure_perm": lambda: mg.pressure_system(12, seed=92, perm_seed=5)}[case]()
    m = s.matrix
    rng = np.random.default_rng(93)
    p64 = rng.standard_normal(m.n)
    p = torch.from_numpy(p64).to(cuda)
    p32 = p.float()
    monkeypatch.setenv("TFLOWS_FP64_OP", "sliced")
    monkeypatch.setenv("TFLOWS_FP32_OP", "sliced")
    b = torch.from_numpy(s.b.copy()).to(cuda)
    out = {}
    for name, env in LAYOUTS.items():
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        ds = pkg.krylov.DeviceSystem(m)
        lay = (ds.fp64_layout(), ds.fp32_layout())
        q = torch.empty_like(p)
        ds.spmv(p, q)
        q32 = torch.empty_like(p32)
        ds.spmv32(p32, q32)
        runs = []
        for method, prec in (("cg", 1), ("cgs", 1), ("cg", 2)):
            x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
            r = b.clone()
            runs.append((ds.solve(method, prec, x, r, 2000, s.tol), x.cpu().numpy(), r.cpu().numpy()))
        xm = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        mres = ds.solve_mixed_reliable(1, xm, b, 3000, s.tol, interval=40)
        torch.cuda.synchronize()
        out[name] = (lay, q.cpu().numpy(), q32.cpu().numpy(), runs, mres, xm.cpu().numpy())
        ds.close()
    shared = out["shared"]
    print(f"{case}: layouts {shared[0]}", flush=True)
    if case == "pressure_perm":
        assert not shared[0][0]["shared"] and not shared[0][1]["shared"]
    else:
        assert shared[0][0]["shared"] and shared[0][1]["shared"]
        assert shared[0][0]["symmetric"] and shared[0][1]["symmetric"]
        assert not out["shared_full"][0][0]["symmetric"] and not out["shared_full"][0][1]["symmetric"]
        # symmetric: about (nnz + N) / 2 FP64 slots and (nnz - N) / 2 FP32 ones
        full = out["shared_full"][0]
        assert full[0]["stored_entries"] <= 1.15 * m.nnz
        assert shared[0][0]["stored_entries"] <= 0.6 * full[0]["stored_entries"]
        assert shared[0][1]["stored_entries"] <= 0.55 * full[1]["stored_entries"]
    assert np.array_equal(shared[1], oracle.spmv(m, p64))
    for other in ("int32", "col16", "shared_full", "shared_full_unroll2", "shared_unroll2"):
        o = out[other]
        assert np.array_equal(o[1], shared[1]) and np.array_equal(o[2], shared[2]), other
        for (ra, xa, rra), (rb, xb, rrb) in zip(o[3], shared[3]):
            assert ra == rb and np.array_equal(xa, xb) and np.array_equal(rra, rrb), other
        assert o[4][:3] == shared[4][:3] and np.array_equal(o[5], shared[5]), other
    assert shared[3][0][0][2] < s.tol and shared[4][2] < s.tol


def test_symmetric_slots_fall_back(pkg, oracle, cuda, monkeypatch):
    """A symmetric pattern whose values are not symmetric (the momentum
    matrix; a pressure matrix given such values later) keeps the full
    shared-offset form: the device compares every lower entry with its
    mirror whenever the values change.  The results are the full form's
    bits, and the SpMVs the oracle's."""
    import torch
    mg = pkg.meshgen
    monkeypatch.setenv("TFLOWS_FP64_OP", "sliced")
    monkeypatch.setenv("TFLOWS_FP32_OP", "sliced")
    mom = mg.momentum_system(14, seed=94).matrix
    pre = mg.pressure_system(14, seed=95)
    rng = np.random.default_rng(96)
    asym = pre.matrix.aval.copy()
    off = np.flatnonzero(pre.matrix.row_of_entry() != pre.matrix.arow[:pre.matrix.nnz] - 1)
    asym[off] *= 1.0 + 0.01 * rng.random(off.size)
    p64 = rng.standard_normal(pre.matrix.n)

    def run(m, vals, sym):
        monkeypatch.setenv("TFLOWS_SYM_SLOTS", sym)
        monkeypatch.setenv("TFLOWS_SYM_SLOTS32", sym)
        ds = pkg.krylov.DeviceSystem(m)
        lay0 = ds.fp64_layout()
        if vals is not None:
            ds.set_values(torch.from_numpy(vals).to(cuda))
        p = torch.from_numpy(p64).to(cuda)
        q, q32 = torch.empty_like(p), torch.empty_like(p.float())
        ds.spmv(p, q)
        ds.spmv32(p.float(), q32)
        b = torch.from_numpy(pre.b.copy()).to(cuda)
        x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        r = b.clone()
        res = ds.solve("cgs", 1, x, r, 300, pre.tol)
        out = (lay0, ds.fp64_layout(), ds.fp32_layout(), q.cpu().numpy(), q32.cpu().numpy(), res,
               x.cpu().numpy(), r.cpu().numpy())
        ds.close()
        return out

    for m, vals in ((mom, None), (pre.matrix, asym)):
        a, b = run(m, vals, "1"), run(m, vals, "0")
        print(f"n {m.n}: layouts {a[0]} -> {a[1]}, {a[2]} / {b[1]}, {b[2]}", flush=True)
        assert a[1] == b[1] and a[2] == b[2]          # the full form in both
        assert not a[1]["symmetric"] and not a[2]["symmetric"]
        for u, v in zip(a[3:], b[3:]):
            assert (u == v) if not isinstance(u, np.ndarray) else np.array_equal(u, v)
        mv = m if vals is None else m.with_values(vals)
        assert np.array_equal(a[3], oracle.spmv(mv, p64))
    # the pressure system had started out symmetric
    sym0 = run(pre.matrix, None, "1")[0]
    assert sym0["symmetric"] and sym0["stored_entries"] < a[1]["stored_entries"]
