"""GPU tests of the pressure-correction assembly (Process/CalcPS.f90:51-202).

Against the oracle (oracle/calcps_oracle.c) on the same inputs: Aval, Flux,
Abou, errmax and b before the mean are bitwise (the diagonal and b of each
cell are summed over its faces in ascending face order, as the reference's
face loop reaches them); the mean bc_sum/vol is a tree sum on the GPU and a
sequential one in the reference, so it agrees to the rounding of the sum,
and b equals the oracle's b-before-the-mean minus the GPU's mean bit for bit.
End to end, the assembled system solved on the GPU matches the oracle's
assembly solved by the oracle's CG.
"""
import math

import numpy as np
import pytest

from test_calcps_oracle import _two_cell_case

pytestmark = pytest.mark.gpu

ALL_TYPES = {"xmin": "inflow", "xmax": "outflow", "ymin": "wall", "ymax": "convect", "zmin": "pressure",
             "zmax": "buffer"}


def _cases(pkg):
    mg = pkg.meshgen
    return {
        "two_cell": _two_cell_case(pkg),
        "channel": mg.calcps_case((12, 10, 8), seed=31),
        "all_types_shuffled": mg.calcps_case((11, 9, 7), seed=32, bc=ALL_TYPES, face_seed=4),
        "nmat2": mg.calcps_case((7, 6, 5), seed=33, bc=ALL_TYPES, nmat=2),
    }


def _exact_mean(case, b_pre):
    """bc_sum / vol with the sums correctly rounded (math.fsum).  b_pre is b
    after the last time-term pass; pass m of Nmat summed b_pre + (Nmat - m)
    times the time term."""
    V = case.volume[case.nbc + 1:]
    term = (V / case.dt) * ((1.5 * case.rho_n[case.nbc + 1:] - 2.0 * case.rho_o[case.nbc + 1:]) +
                            0.5 * case.rho_oo[case.nbc + 1:])
    parts = [(b_pre + (case.nmat - m) * term) * V for m in range(1, case.nmat + 1)]
    vols = [V] * case.nmat
    return math.fsum(np.concatenate(parts)) / math.fsum(np.concatenate(vols)), np.abs(np.concatenate(parts)).sum()


@pytest.mark.parametrize("name", ["two_cell", "channel", "all_types_shuffled", "nmat2"])
def test_assembly_matches_oracle(pkg, oracle, cuda, name):
    case = _cases(pkg)[name]
    want = oracle.calcps(case)
    dc = pkg.krylov.DeviceCalcPS(case)
    got = dc.assemble()
    aval = got["aval"].cpu().numpy()
    assert np.array_equal(aval, want["aval"])
    assert np.array_equal(got["flux"].cpu().numpy()[:case.ns], want["flux"])
    assert np.array_equal(got["abou"].cpu().numpy()[:case.nbc], want["abou"])
    assert got["errmax"] == want["errmax"]
    # the mean: GPU tree sum vs the correctly rounded sum, and vs the
    # reference's sequential sum, within the rounding of a sum of
    # NC * Nmat terms
    exact, mag = _exact_mean(case, want["b_pre"])
    vol = case.volume[case.nbc + 1:].sum() * case.nmat
    bound = 4 * case.nc * case.nmat * np.finfo(float).eps * mag / vol
    assert abs(got["mean"] - exact) <= bound + 1e-300
    assert abs(want["mean"] - exact) <= bound + 1e-300
    b = got["b"].cpu().numpy()
    assert np.array_equal(b, want["b_pre"] - got["mean"])
    # the same assembly from host arrays, and a repeat, give the same bits
    host = dc.assemble_host()
    assert np.array_equal(host["aval"], aval) and np.array_equal(host["b"], b)
    again = dc.assemble()
    assert np.array_equal(again["b"].cpu().numpy(), b) and again["mean"] == got["mean"]
    dc.close()


def test_nmat_zero_keeps_flux(pkg, oracle, cuda):
    """Nmat = 0 runs no face pass: Aval is zero, Flux and Abou keep what the
    caller left there, b = 0 - 0/0 (NaN), as in the reference."""
    import torch
    case = _cases(pkg)["two_cell"]
    dc = pkg.krylov.DeviceCalcPS(case)
    out = dict(aval=torch.full((case.nonz,), 7.0, dtype=torch.float64, device=cuda),
               b=torch.empty(case.nc, dtype=torch.float64, device=cuda),
               flux=torch.full((case.ns,), 3.0, dtype=torch.float64, device=cuda),
               abou=torch.full((case.nbc,), 5.0, dtype=torch.float64, device=cuda))
    got = dc.assemble(nmat=0, out=out)
    want = oracle.calcps(case, nmat=0)
    assert np.array_equal(got["aval"].cpu().numpy(), want["aval"])
    assert np.all(got["flux"].cpu().numpy() == 3.0) and np.all(got["abou"].cpu().numpy() == 5.0)
    assert got["errmax"] == 0.0 and np.all(np.isnan(got["b"].cpu().numpy()))
    dc.close()


@pytest.mark.parametrize("prec", [1, 2])
def test_assembled_pressure_solve(pkg, oracle, cuda, prec):
    """GPU assembly into the device system, GPU CG, against the oracle's
    assembly solved by the oracle's CG (CalcPS.f90:214-219: x0 = 0, cap
    800, absolute tolerance)."""
    import torch
    case = pkg.meshgen.calcps_case((14, 12, 10), seed=34)
    want = oracle.calcps(case)
    m = case.matrix(want["aval"])
    tol = 1e-8 * np.linalg.norm(want["b"])
    ref = oracle.run("cg", m, want["b"], prec=prec, niter=800, tol=tol)
    dc = pkg.krylov.DeviceCalcPS(case)
    got = dc.assemble()
    ds = pkg.krylov.DeviceSystem(m)
    ds.set_values(got["aval"])
    x = torch.zeros(case.nc, dtype=torch.float64, device=cuda)
    r = got["b"].clone()
    it, ini, fin = ds.solve("cg", prec, x, r, 800, tol)
    torch.cuda.synchronize()
    assert abs(it - ref["niter"]) <= 1 and fin < tol
    A = m.to_scipy()
    xs = x.cpu().numpy()
    # both residuals are below tol; x agrees within what that fixes
    assert np.linalg.norm(want["b"] - A @ xs) < 1.01 * tol
    inv = 1.0 / np.linalg.svd(A.toarray(), compute_uv=False)[-1]
    assert np.linalg.norm(xs - ref["x"]) <= 1.01 * inv * (fin + ref["fin_res"]) + 1e-12 * np.linalg.norm(ref["x"])
    ds.close()
    dc.close()


@pytest.mark.parametrize("name", ["two_cell", "channel", "all_types_shuffled"])
def test_correction_matches_oracle(pkg, oracle, cuda, name):
    """CorUVW.f90:36-121 on the GPU: corrected velocities (PRESSURE
    boundary copies included), corrected fluxes, the mass error per unit
    volume and its maximum are the oracle's bit for bit."""
    import torch
    case = _cases(pkg)[name]
    asm = oracle.calcps(case)
    rng = np.random.default_rng(23)
    pp = rng.standard_normal(case.nbc + case.nc + 1) * 1e-2
    want = oracle.coruvw(case, pp, asm["aval"], asm["abou"], asm["flux"])
    dc = pkg.krylov.DeviceCalcPS(case)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(cuda)   # noqa: E731
    u, v, w = t(case.u), t(case.v), t(case.w)
    flux = t(asm["flux"])
    abou = t(asm["abou"]) if case.nbc else torch.zeros(1, dtype=torch.float64, device=cuda)
    b, em = dc.correct(t(pp), t(asm["aval"]), abou, u, v, w, flux)
    torch.cuda.synchronize()
    for got, key in ((u, "u"), (v, "v"), (w, "w"), (flux, "flux"), (b, "b")):
        assert np.array_equal(got.cpu().numpy(), want[key]), key
    assert em == want["errmax"]
    dc.close()
