"""GPU tests of the four forms of the DIC sweeps.

The DIC apply (Process/CG.f90:135-156) runs over the matrix rows level by
level (TFLOWS_DIC_SWEEP=0), over a level-major copy of the triangles level
by level (1: rows of one forward level at consecutive positions, entries in
matrix order), over that copy in one launch per triangle with each row
waiting for its dependencies (2), or in one launch per triangle over compact
tiles of up to 8 levels, each swept by one wavefront with its values in LDS
(3, the default where the pattern allows it; the pivots of Prec1.f90:51-59
then take one launch too).  Every z_i and every pivot is the same sum of the
same products in all four, so solves must agree bit for bit: single systems
in natural and locality order, the three methods, FP32 inner solves,
batches, and after a value update.  The form is read when a system first
builds its levels; a pattern that does not allow tiles (a randomly
renumbered one has rows with more than 4 entries per triangle) runs form 2
when 3 is asked for.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


FORMS = (0, 1, 2, 3)


def _forms(monkeypatch, make):
    """One object per sweep form; `make(form)` creates it and runs its
    first DIC use.  Returns the objects."""
    out = []
    for f in FORMS:
        monkeypatch.setenv("TFLOWS_DIC_SWEEP", str(f))
        out.append(make(f))
    monkeypatch.delenv("TFLOWS_DIC_SWEEP")
    return out


def _system(pkg, case):
    mg = pkg.meshgen
    if case == "natural":
        return mg.pressure_system(14, seed=61, prec=2)
    if case == "renumbered":
        return mg.pressure_system(12, seed=62, perm_seed=7, prec=2)
    return mg.momentum_system(11, seed=63, prec=2)


@pytest.mark.parametrize("case", ["natural", "renumbered", "momentum"])
@pytest.mark.parametrize("method", ["cg", "bicg", "cgs"])
def test_sweep_forms_solves_bitwise(pkg, cuda, monkeypatch, case, method):
    import torch
    s = _system(pkg, case)
    m = s.matrix
    runs = {}

    def make(form):
        ds = pkg.krylov.DeviceSystem(m)
        out = []
        for cap in (7, 400):
            x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
            r = torch.from_numpy(s.b.copy()).to(cuda)
            it, ini, fin = ds.solve(method, 2, x, r, cap, s.tol)
            torch.cuda.synchronize()
            out.append((it, ini, fin, x.cpu().numpy(), r.cpu().numpy()))
        runs[form] = out
        return ds

    objs = _forms(monkeypatch, make)
    for f in FORMS[1:]:
        for a, b in zip(runs[0], runs[f]):
            assert a[:3] == b[:3], f
            assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4]), f
    if case != "momentum" or method == "cgs":
        assert runs[3][1][2] < s.tol
    for o in objs:
        o.close()


@pytest.mark.parametrize("method", ["cg", "cgs"])
def test_sweep_forms_after_value_update(pkg, cuda, monkeypatch, method):
    """New values reach the level-major copy before the next sweep."""
    import torch
    s = _system(pkg, "momentum")
    m = s.matrix
    scale = torch.from_numpy(1.0 + 0.25 * np.random.default_rng(3).random(m.nonz)).to(cuda)
    runs = {}

    def make(form):
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
        runs[form] = out
        return ds

    objs = _forms(monkeypatch, make)
    for f in FORMS[1:]:
        for a, b in zip(runs[0], runs[f]):
            assert a[:2] == b[:2] and np.array_equal(a[2], b[2]), f
    assert not np.array_equal(runs[3][0][2], runs[3][1][2])
    for o in objs:
        o.close()


@pytest.mark.parametrize("method", ["cg", "cgs"])
def test_sweep_forms_mixed_bitwise(pkg, cuda, monkeypatch, method):
    """FP32 inner solves sweep the FP32 copy, rounded from the same FP64
    values as the FP32 matrix."""
    import torch
    s = _system(pkg, "momentum" if method == "cgs" else "renumbered")
    m = s.matrix
    runs = {}

    def make(form):
        ds = pkg.krylov.DeviceSystem(m)
        x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        b = torch.from_numpy(s.b.copy()).to(cuda)
        res = ds.solve_mixed(method, 2, x, b, 2000, s.tol, inner_rtol=1e-3)
        torch.cuda.synchronize()
        runs[form] = (res, x.cpu().numpy())
        return ds

    objs = _forms(monkeypatch, make)
    for f in FORMS[1:]:
        assert runs[0][0] == runs[f][0], f
        assert np.array_equal(runs[0][1], runs[f][1]), f
    assert runs[3][0][2] < s.tol
    for o in objs:
        o.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_sweep_forms_batch_bitwise(pkg, cuda, monkeypatch, mixed):
    """Batches sweep the whole batch's levels; fixed grid and lanes so that
    both runs use the same reduction order."""
    import torch
    systems = pkg.meshgen.transport_batch(n=10, count=4, seed=5, prec=2)
    runs = {}

    def make(form):
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
        runs[form] = (res, x.cpu().numpy(), r.cpu().numpy())
        return B

    objs = _forms(monkeypatch, make)
    for f in FORMS[1:]:
        for key in ("niter", "fin_res"):
            assert np.array_equal(runs[0][0][key], runs[f][0][key]), (f, key)
        assert np.array_equal(runs[0][1], runs[f][1]), f
        assert np.array_equal(runs[0][2], runs[f][2]), f
    assert np.all(runs[3][0]["fin_res"] < np.array([t.tol for t in systems]))
    for o in objs:
        o.close()


@pytest.mark.parametrize("form, latched", [(2, 1), (3, 0)])
def test_sweep_wait_bound(pkg, cuda, monkeypatch, form, latched):
    """Failure handling of the one-launch sweeps, forced from the host: with
    TFLOWS_DIC_WAIT_MS=0 (read when the levels are built) a row that has to
    wait gives up at its first time check, raises the fault word and leaves
    NaN; the other waiting rows see the fault and give up too, so every
    wavefront ends.  The solve then restores x and r1, latches the system to
    one launch per level (form 1 after form 2, form 0 after form 3; neither
    needs a co-resident grid) and runs again, so it returns the unbounded
    result bit for bit -- as it does when no row had to wait.  Later solves
    on that system stay in the latched form and keep the same bits.  A
    system built afterwards with the default bound solves normally in the
    one-launch form."""
    import torch
    s = pkg.meshgen.pressure_system(48, seed=66, prec=2)
    m = s.matrix
    monkeypatch.setenv("TFLOWS_DIC_SWEEP", str(form))

    def run(ds, cap):
        x = torch.zeros(m.n, dtype=torch.float64, device=cuda)
        r = torch.from_numpy(s.b.copy()).to(cuda)
        res = ds.solve("cg", 2, x, r, cap, s.tol)
        torch.cuda.synchronize()
        return res, x.cpu().numpy(), r.cpu().numpy()

    ref = pkg.krylov.DeviceSystem(m)
    want = run(ref, 20)
    assert ref.dic_sweep() == form
    monkeypatch.setenv("TFLOWS_DIC_WAIT_MS", "0")
    ds = pkg.krylov.DeviceSystem(m)
    got = run(ds, 20)
    now = ds.dic_sweep()
    print(f"wait bound 0: sweep form after the solve = {now}", flush=True)
    assert now in (latched, form)
    assert got[0] == want[0] and np.array_equal(got[1], want[1]) and np.array_equal(got[2], want[2])
    again_same = run(ds, 20)
    assert ds.dic_sweep() == now
    assert again_same[0] == want[0] and np.array_equal(again_same[1], want[1])
    ds.close()
    monkeypatch.delenv("TFLOWS_DIC_WAIT_MS")
    again = pkg.krylov.DeviceSystem(m)
    res, x, _ = run(again, 20)
    assert again.dic_sweep() == form
    assert res == want[0] and np.array_equal(x, want[1])
    again.close()
    ref.close()


def test_tiles_taken_where_the_pattern_allows(pkg, cuda, monkeypatch):
    """The default form: tiles on naturally numbered boxes (locality order
    inside the library or not); row-by-row waits on a randomly renumbered
    pattern, whose rows have up to 6 entries per triangle."""
    monkeypatch.delenv("TFLOWS_DIC_SWEEP", raising=False)
    mg = pkg.meshgen
    for sysm, want in ((mg.pressure_system(20, seed=1, prec=2), 3),
                       (mg.pressure_system(12, seed=62, perm_seed=7, prec=2), 2)):
        ds = pkg.krylov.DeviceSystem(sysm.matrix)
        assert ds.dic_sweep() == want, want
        ds.close()
