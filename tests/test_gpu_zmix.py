"""GPU tests of the mixture-fraction assembly (Process/CalcZmix.f90:100-405).

Against the oracle (oracle_calczmix, oracle_zmix_update): Aval, b, Abou and
rhoZ bit for bit on the cases and flag sets of the momentum tests, from
device and from host arrays; the limited update bit for bit; end to end, the
assembled system solved with the GPU CG (cap 20 as in SIMPLE, :388-394)
against the oracle's assembly and CG.
"""
import numpy as np
import pytest

from test_gpu_newuvw import CASES, SCHEMES, _case, with_mass_flux
from test_zmix_oracle import _hand_zfields

pytestmark = pytest.mark.gpu

YES, NO, CN, LIN, PAR = 10, 11, 2, 4, 5


def _zcase(pkg, name):
    case, uf = _case(pkg, name)
    if name == "two_cell":
        zf = _hand_zfields(case, uf["flux_u"])
        for k, g in zip(("zx", "zy", "zz"), (0.5, -1.0, 0.25)):
            zf[k] = np.full_like(zf["z"], g)
        return case, zf
    zf = pkg.meshgen.zmix_fields(case, seed=71, nmaterials=uf["nmaterials"])
    zf.update({k: uf[k] for k in ("scoef", "ff", "material", "flux_u")})
    return case, zf


def _zparams(pkg, name, nmat):
    sets = {
        "implicit": dict(),
        "upwind_par_cn": dict(blend_tem=[YES] * nmat, inert=PAR, diffus=CN, urf=0.6, urfc=[0.9, 0.7][:nmat]),
        "upwind_part_only": dict(blend_tem=[YES] * nmat, convec=1, cross=1, urf=1.0),
        "no_diffusion": dict(diffus=1, blend_tem=[YES] * nmat, inert=0, visc=0.05),
        "density_weighted_visc": dict(visc_rho=1, blend_tem=[YES] * nmat, visc=0.03),
    }
    return pkg.meshgen.uvw_params(nmaterials=nmat, **sets[name])


@pytest.mark.parametrize("flags", ["implicit", "upwind_par_cn", "upwind_part_only", "no_diffusion",
                                   "density_weighted_visc"])
@pytest.mark.parametrize("name", CASES)
def test_assembly_bitwise(pkg, oracle, cuda, name, flags):
    case, zf = _zcase(pkg, name)
    p = _zparams(pkg, flags, zf["nmaterials"])
    want = oracle.calczmix(case, zf, p)
    dc = pkg.krylov.DeviceCalcPS(case)
    dc.uvw_setup(zf)
    args = pkg.meshgen.zmix_args(case, zf)
    print(f"case {name} flags {flags} nc {case.nc} ns {case.ns}", flush=True)
    got = {k: v.cpu().numpy() for k, v in dc.assemble_zmix(p, dc.zmix_inputs(args)).items()}
    assert np.array_equal(got["aval"], want["aval"])
    assert np.array_equal(got["b"], want["b"])
    assert np.array_equal(got["abou"][:case.nbc], want["abou"])
    assert np.array_equal(got["rhoz"], want["rhoz"])
    host = dc.assemble_zmix_host(p, args)
    for k in ("aval", "b", "rhoz"):
        assert np.array_equal(host[k], got[k]), k
    dc.close()


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("name", CASES)
def test_conv_scheme_bitwise(pkg, oracle, cuda, name, scheme):
    """ConvScheme through the scalar assembly (BLEND_TEM, with rhoZ and
    dZ/dx, dZ/dy, dZ/dz): bit for bit the oracle's, every scheme, a mass
    flux differing in sign from Flux_u on some faces."""
    case, zf = _zcase(pkg, name)
    zf = with_mass_flux(zf)
    nmat = zf["nmaterials"]
    p = pkg.meshgen.uvw_params(nmaterials=nmat, blend_tem=[scheme, NO][:nmat], urf=0.7, urfc=[0.9, 0.8][:nmat])
    want = oracle.calczmix(case, zf, p)
    dc = pkg.krylov.DeviceCalcPS(case)
    dc.uvw_setup(zf)
    args = pkg.meshgen.zmix_args(case, zf)
    print(f"case {name} scheme {scheme}", flush=True)
    got = {k: v.cpu().numpy() for k, v in dc.assemble_zmix(p, dc.zmix_inputs(args)).items()}
    for k in ("aval", "b", "rhoz"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["abou"][:case.nbc], want["abou"])
    host = dc.assemble_zmix_host(p, args)
    assert np.array_equal(host["b"], got["b"])
    dc.close()


def test_blend_tem_schemes_refused(pkg, cuda):
    mg = pkg.meshgen
    case, zf = _zcase(pkg, "two_materials")
    dc = pkg.krylov.DeviceCalcPS(case)
    dc.uvw_setup(zf)
    inputs = dc.zmix_inputs(mg.zmix_args(case, zf))
    for bt in ([28, 28], [YES, NO]):
        with pytest.raises(pkg.krylov.TFlowsError):
            dc.assemble_zmix(mg.uvw_params(nmaterials=2, blend_tem=bt), inputs)
    dc.assemble_zmix(mg.uvw_params(nmaterials=2, blend_tem=[23, 23]), inputs)
    # BLEND (momentum) is not read by the scalar assembly
    dc.assemble_zmix(mg.uvw_params(nmaterials=2, blend=[21, 21], blend_tem=[NO, NO]), inputs)
    dc.close()


def test_assemble_solve_update(pkg, oracle, cuda):
    """CalcZmix end to end: GPU assembly, GPU CG from rhoZ (cap 20,
    absolute tolerance), Zmix%n = rhoZ / rho limited to [0, 1]; against
    the oracle's assembly, CG and update.  The update is bitwise given the
    same rhoZ; the solutions agree within ||A^-1|| (r_gpu + r_ref)."""
    import torch
    mg = pkg.meshgen
    case = mg.calcps_case((8, 6, 7), seed=72)
    zf = mg.zmix_fields(case, seed=73)
    dc = pkg.krylov.DeviceCalcPS(case)
    dc.uvw_setup(zf)
    nb = case.nbc
    p = mg.uvw_params(blend_tem=[YES], urf=0.8, visc=2e-2)
    want = oracle.calczmix(case, zf, p)
    m = case.matrix(want["aval"])
    tol = 1e-12 * np.linalg.norm(want["b"])
    ref = oracle.run("cg", m, want["b"], x0=want["rhoz"][nb + 1:], prec=1, niter=20, tol=tol)
    args = mg.zmix_args(case, zf)
    out = dc.assemble_zmix(p, dc.zmix_inputs(args))
    ds = pkg.krylov.DeviceSystem(m)
    ds.set_values(out["aval"])
    x = out["rhoz"][nb + 1:].clone()
    r = out["b"].clone()
    it, ini, fin = ds.solve("cg", 1, x, r, 20, tol)
    out["rhoz"][nb + 1:] = x
    z = torch.from_numpy(np.ascontiguousarray(args["z_n"])).cuda()
    rho = torch.from_numpy(case.rho_n).cuda()
    dc.zmix_update(out["rhoz"], rho, z)
    torch.cuda.synchronize()
    assert abs(it - ref["niter"]) <= 1 and fin < tol
    A = m.to_scipy().toarray()
    inv = 1.0 / np.linalg.svd(A, compute_uv=False)[-1]
    xs = x.cpu().numpy()
    assert np.linalg.norm(xs - ref["x"]) <= 1.01 * inv * (fin + ref["fin_res"]) + 1e-14 * np.linalg.norm(ref["x"])
    rz = out["rhoz"].cpu().numpy()
    got = z.cpu().numpy()
    assert np.array_equal(got, oracle.zmix_update(case, rz, case.rho_n, args["z_n"]))
    assert np.all((got[nb + 1:] >= 0.0) & (got[nb + 1:] <= 1.0))
    ds.close()
    dc.close()


@pytest.mark.parametrize("flags,scheme", [("implicit", None), ("upwind_par_cn", None), ("implicit", 27)])
def test_assembly_bitwise_block_order(pkg, oracle, cuda, flags, scheme):
    """The scalar form with the block-shaped visiting order (64 x 64 x 40
    channel; three gradients per cell in LDS): Aval, b, Abou and rhoZ bit for
    bit the oracle's, with and without a face scheme."""
    mg = pkg.meshgen
    case = mg.calcps_case((64, 64, 40), seed=73)
    uf = with_mass_flux(mg.uvw_fields(case, seed=74))
    zf = mg.zmix_fields(case, seed=75, nmaterials=1)
    zf.update({k: uf[k] for k in ("scoef", "ff", "material", "flux_u", "flux")})
    p = _zparams(pkg, flags, 1)
    if scheme is not None:
        p = dict(p, blend_tem=[scheme])
    want = oracle.calczmix(case, zf, p)
    dc = pkg.krylov.DeviceCalcPS(case)
    dc.uvw_setup(zf)
    args = mg.zmix_args(case, zf)
    print(f"visit order {dc.visit_order()} flags {flags} scheme {scheme}", flush=True)
    assert dc.visit_order() == 2
    got = {k: v.cpu().numpy() for k, v in dc.assemble_zmix(p, dc.zmix_inputs(args)).items()}
    for k in ("aval", "b", "rhoz"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["abou"][:case.nbc], want["abou"])
    dc.close()

