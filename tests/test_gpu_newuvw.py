"""GPU tests of the momentum assembly (Process/NewUVW.f90:137-499).

Against the oracle (oracle/calcps_oracle.c, oracle_newuvw) on the same
inputs: Aval, b, Abou, Asave and rhoUi are the oracle's bit for bit -- the
GPU sums the two face loops of every cell over the cell's faces in ascending
face order, the order in which the reference's loops reach the cell, and
computes each face's terms in the reference's operation order.  Cases: the
two-cell mesh with every boundary type, channel and shuffled meshes, two
materials, faces that share an off-diagonal (split faces), U, V and W, and
the flag sets the assembly covers, and every ConvScheme face scheme.  End
to end, U, V and W are assembled,
solved with the GPU BiCG and updated (:505-526), against the oracle's
assembly solved by the oracle's BiCG.
"""
import dataclasses

import numpy as np
import pytest

from test_newuvw_oracle import CLOSED_BC, _hand_case, _hand_fields

pytestmark = pytest.mark.gpu

YES, NO, FI, CN, LIN, PAR = 10, 11, 3, 2, 4, 5
ALL_TYPES = {"xmin": "inflow", "xmax": "outflow", "ymin": "wall", "ymax": "convect", "zmin": "pressure",
             "zmax": "buffer"}


def _split_faces(mg, case, every=5):
    """Every ``every``-th interior face split into two halves (half the area,
    same centre distance, same weight); both halves name the same
    off-diagonal positions, so several faces write one position."""
    inner = np.nonzero(case.sidec[2] > 0)[0][::every]
    order = np.concatenate([np.arange(case.ns), inner])
    half = np.ones(len(order))
    half[inner] = 0.5
    half[case.ns:] = 0.5
    pick = lambda a: np.ascontiguousarray(a[..., order])                # noqa: E731
    return dataclasses.replace(case, ns=len(order), sidec=pick(case.sidec), sidaij=pick(case.sidaij),
                               sx=pick(case.sx) * half, sy=pick(case.sy) * half, sz=pick(case.sz) * half,
                               dx=pick(case.dx), dy=pick(case.dy), dz=pick(case.dz), f=pick(case.f))


def _case(pkg, name):
    mg = pkg.meshgen
    if name == "two_cell":
        case = _hand_case(pkg)
        return case, _hand_fields(case, [3.0, 1.0, -2.0, 0.5, 0.0, -1.0],
                                  {"ux": 1.0, "uy": 2.0, "vx": 3.0, "wz": -1.0, "vy": 0.5, "wx": -0.25})
    if name == "channel":
        case = mg.calcps_case((12, 10, 8), seed=51)
        return case, mg.uvw_fields(case, seed=52)
    if name == "all_types_shuffled":
        case = mg.calcps_case((11, 9, 7), seed=53, bc=ALL_TYPES, face_seed=6)
        return case, mg.uvw_fields(case, seed=54, zero_flux_types=("wall",))
    if name == "two_materials":
        case = mg.calcps_case((10, 8, 6), seed=55, bc=CLOSED_BC)
        return case, mg.uvw_fields(case, seed=56, nmaterials=2)
    if name == "split_faces":
        case = _split_faces(mg, mg.calcps_case((9, 7, 6), seed=57, bc=ALL_TYPES))
        return case, mg.uvw_fields(case, seed=58, zero_flux_types=())
    raise KeyError(name)


def _params(pkg, name, nmat, var):
    mg = pkg.meshgen
    sets = {
        "implicit": dict(),
        "upwind_par_cn": dict(blend=[YES] * nmat, blend_tem=[YES] * nmat, inert=PAR, diffus=CN, urf=0.6,
                              urfc=[0.9, 0.7][:nmat]),
        "upwind_part_only": dict(blend_tem=[YES] * nmat, convec=1, cross=1, urf=1.0),
        "no_diffusion": dict(diffus=1, blend=[YES] * nmat, inert=0, visc=0.05),
        "density_weighted_visc": dict(visc_rho=1, blend=[YES] * nmat, blend_tem=[YES] * nmat, visc=0.03),
    }
    return mg.uvw_params(nmaterials=nmat, var=var, **sets[name])


CASES = ["two_cell", "channel", "all_types_shuffled", "two_materials", "split_faces"]
FLAGS = ["implicit", "upwind_par_cn", "upwind_part_only", "no_diffusion", "density_weighted_visc"]


def _run(pkg, case, uf, p):
    dc = pkg.krylov.DeviceCalcPS(case)
    dc.uvw_setup(uf)
    args = pkg.meshgen.uvw_args(case, uf, p["var"])
    out = dc.assemble_uvw(p, dc.uvw_inputs(args))
    return dc, args, {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("var", [1, 2, 3])
@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("name", CASES)
def test_assembly_bitwise(pkg, oracle, cuda, name, flags, var):
    case, uf = _case(pkg, name)
    p = _params(pkg, flags, uf["nmaterials"], var)
    want = oracle.newuvw(case, uf, p)
    dc, args, got = _run(pkg, case, uf, p)
    print(f"case {name} flags {flags} var {var} nc {case.nc} ns {case.ns}", flush=True)
    assert np.array_equal(got["aval"], want["aval"])
    assert np.array_equal(got["b"], want["b"])
    assert np.array_equal(got["abou"][:case.nbc], want["abou"])
    assert np.array_equal(got["rhoui"], want["rhoui"])
    nb = case.nbc
    if var == 1:
        assert np.array_equal(got["asave"][nb + 1:], want["asave"][nb + 1:])
    else:
        assert np.all(got["asave"] == 0.0)
    # host arrays in and out give the same bits
    host = dc.assemble_uvw_host(p, args)
    for k in ("aval", "b", "rhoui"):
        assert np.array_equal(host[k], got[k]), k
    dc.close()


def test_shared_offdiagonals_present(pkg):
    """The split-face case does write one off-diagonal from two faces."""
    case, _ = _case(pkg, "split_faces")
    inner = case.sidec[2] > 0
    pos = case.sidaij[0][inner]
    assert len(np.unique(pos)) < len(pos)


SCHEMES = list(range(20, 28))   # CDS .. GAMMA (Process/IniPar.f90:42-49)


def with_mass_flux(uf):
    """The mass flux Flux that picks ConvScheme's upwind cell: Flux_u with
    every third face's sign flipped, so there the face scheme's upwind side
    and the upwind part's differ."""
    mf = np.array(uf["flux_u"], dtype=float)
    mf[::3] *= -1.0
    return dict(uf, flux=mf)


@pytest.mark.parametrize("var", [1, 2, 3])
@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("name", CASES)
def test_conv_scheme_bitwise(pkg, oracle, cuda, name, scheme, var):
    """ConvScheme's face values (with CalMinMax's bounds) on the GPU: Aval,
    b, Abou, rhoUi bit for bit the oracle's, for every scheme, U, V and W,
    with a mass flux whose sign differs from Flux_u's on some faces; on two
    materials the scheme sits next to NO."""
    case, uf = _case(pkg, name)
    uf = with_mass_flux(uf)
    nmat = uf["nmaterials"]
    p = pkg.meshgen.uvw_params(nmaterials=nmat, var=var, blend=[scheme, NO][:nmat], blend_tem=[YES] * nmat,
                               urf=0.7, urfc=[0.9, 0.8][:nmat])
    want = oracle.newuvw(case, uf, p)
    dc, args, got = _run(pkg, case, uf, p)
    print(f"case {name} scheme {scheme} var {var}", flush=True)
    assert np.array_equal(got["aval"], want["aval"])
    assert np.array_equal(got["b"], want["b"])
    assert np.array_equal(got["abou"][:case.nbc], want["abou"])
    assert np.array_equal(got["rhoui"], want["rhoui"])
    if var == 1:
        host = dc.assemble_uvw_host(p, args)
        for k in ("aval", "b", "rhoui"):
            assert np.array_equal(host[k], got[k]), k
        # without a mass flux Flux_u picks the upwind side, as in the oracle
        uf0 = dict(uf)
        del uf0["flux"]
        want0 = oracle.newuvw(case, uf0, p)
        got0 = dc.assemble_uvw(p, dc.uvw_inputs(pkg.meshgen.uvw_args(case, uf0, 1)))
        assert np.array_equal(got0["b"].cpu().numpy(), want0["b"])
    dc.close()


def test_blend_schemes_refused(pkg, cuda):
    """YES next to NO at a face (the reference leaves the face value
    undefined there), or a code that is no scheme, is refused with an error
    before anything runs; a ConvScheme scheme is not."""
    mg = pkg.meshgen
    case, uf = _case(pkg, "two_materials")
    dc = pkg.krylov.DeviceCalcPS(case)
    dc.uvw_setup(uf)
    inputs = dc.uvw_inputs(mg.uvw_args(case, uf, 1))
    for blend in ([YES, NO], [28, 28], [19, NO]):
        with pytest.raises(pkg.krylov.TFlowsError):
            dc.assemble_uvw(mg.uvw_params(nmaterials=2, blend=blend), inputs)
    dc.assemble_uvw(mg.uvw_params(nmaterials=2, blend=[YES, YES]), inputs)
    dc.assemble_uvw(mg.uvw_params(nmaterials=2, blend=[21, 21]), inputs)
    dc.close()


def test_assemble_solve_update(pkg, oracle, cuda):
    """NewUVW end to end for U, V, W: assembly on the GPU, BiCG on the GPU
    from rhoUi (niter 200, absolute tolerance), Ui%n = rhoUi / rho
    (:505-526); against the oracle's assembly, the oracle's BiCG and the same
    update.  Both residuals are below tol, so the solutions agree within
    ||A^-1|| (r_gpu + r_ref)."""
    import torch
    mg = pkg.meshgen
    case = mg.calcps_case((8, 7, 6), seed=59)
    uf = mg.uvw_fields(case, seed=60)
    dc = pkg.krylov.DeviceCalcPS(case)
    dc.uvw_setup(uf)
    nb = case.nbc
    rho = torch.from_numpy(case.rho_n).cuda()
    for var in (1, 2, 3):
        p = mg.uvw_params(var=var, blend=[YES], blend_tem=[YES], urf=0.7)
        want = oracle.newuvw(case, uf, p)
        m = case.matrix(want["aval"])
        tol = 1e-12 * np.linalg.norm(want["b"])
        ref = oracle.run("bicg", m, want["b"], x0=want["rhoui"][nb + 1:], prec=1, niter=200, tol=tol)
        args = mg.uvw_args(case, uf, var)
        out = dc.assemble_uvw(p, dc.uvw_inputs(args))
        ds = pkg.krylov.DeviceSystem(m)
        ds.set_values(out["aval"])
        x = out["rhoui"][nb + 1:].clone()
        r = out["b"].clone()
        it, ini, fin = ds.solve("bicg", 1, x, r, 200, tol)
        out["rhoui"][nb + 1:] = x
        ui = torch.from_numpy(np.ascontiguousarray(args["ui_n"])).cuda()
        dc.uvw_update(out["rhoui"], rho, ui)
        torch.cuda.synchronize()
        assert abs(it - ref["niter"]) <= 1 and fin < tol
        A = m.to_scipy().toarray()
        inv = 1.0 / np.linalg.svd(A, compute_uv=False)[-1]
        xs = x.cpu().numpy()
        assert np.linalg.norm(xs - ref["x"]) <= 1.01 * inv * (fin + ref["fin_res"]) + 1e-14 * np.linalg.norm(ref["x"])
        u_ref = ref["x"] / case.rho_n[nb + 1:]
        got = ui.cpu().numpy()
        assert np.array_equal(got[nb + 1:], xs / case.rho_n[nb + 1:])       # the update, bit for bit
        np.testing.assert_allclose(got[nb + 1:], u_ref, rtol=1e-9)
        assert np.array_equal(got[:nb + 1], args["ui_n"][:nb + 1])          # boundary values untouched
        ds.close()
    dc.close()


@pytest.mark.parametrize("cap,order", [("1", 1), ("2", 2)])
@pytest.mark.parametrize("flags,var", [("implicit", 1), ("upwind_par_cn", 2), ("no_diffusion", 3)])
def test_assembly_bitwise_visit_orders(pkg, oracle, cuda, monkeypatch, flags, var, cap, order):
    """A naturally numbered channel of 64 x 64 x 40 cells: its planes hold 16
    tiles of 256 cells, its lines one wavefront.  Capped at 1 the cell kernel
    visits its tiles brick by brick (8 tile rows x 32 planes, then the
    remaining 8 planes); by default each block is 64 cells x 2 lines x 2
    planes, whose faces inside the block take the other cell's values from
    LDS.  Every output stays the oracle's bit for bit."""
    mg = pkg.meshgen
    case = mg.calcps_case((64, 64, 40), seed=61)
    uf = mg.uvw_fields(case, seed=62)
    inner = case.sidec[2] > 0
    off = np.abs(case.sidec[2][inner] - case.sidec[1][inner])
    vals, counts = np.unique(off, return_counts=True)
    line, plane = sorted(vals[np.argsort(counts)[-3:]])[1:]
    rows = plane // 256
    assert plane % 256 == 0 and case.nc % plane == 0 and rows > 8 and rows % 8 == 0   # the brick order applies
    assert line % 64 == 0 and (plane // line) % 2 == 0 and (case.nc // plane) % 2 == 0   # so do the blocks
    monkeypatch.setenv("TFLOWS_UVW_ORDER", cap)
    p = _params(pkg, flags, uf["nmaterials"], var)
    want = oracle.newuvw(case, uf, p)
    dc, args, got = _run(pkg, case, uf, p)
    print(f"visit order {dc.visit_order()} (cap {cap}): nc {case.nc} line {line} plane {plane} flags {flags} "
          f"var {var}", flush=True)
    assert dc.visit_order() == order
    for k in ("aval", "b", "rhoui"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["abou"][:case.nbc], want["abou"])
    dc.close()


@pytest.mark.parametrize("scheme,var,visc_rho", [(24, 1, 0), (27, 2, 0), (21, 3, 1)])
def test_conv_scheme_bitwise_block_order(pkg, oracle, cuda, scheme, var, visc_rho):
    """ConvScheme face values with the block-shaped visiting order (64 x 64
    x 40 channel): the upwind extrapolation reads the other cell's gradients
    while its density and velocity come from LDS; bit for bit the oracle's."""
    mg = pkg.meshgen
    case = mg.calcps_case((64, 64, 40), seed=63)
    uf = with_mass_flux(mg.uvw_fields(case, seed=64))
    p = mg.uvw_params(nmaterials=1, var=var, blend=[scheme], blend_tem=[YES], urf=0.7, urfc=[0.9],
                      visc_rho=visc_rho)
    want = oracle.newuvw(case, uf, p)
    dc, args, got = _run(pkg, case, uf, p)
    print(f"visit order {dc.visit_order()} scheme {scheme} var {var} visc_rho {visc_rho}", flush=True)
    assert dc.visit_order() == 2
    for k in ("aval", "b", "rhoui"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["abou"][:case.nbc], want["abou"])
    dc.close()



@pytest.mark.parametrize("name", CASES + ["channel_64x64x40"])
def test_cfl_peclet_monitor(pkg, oracle, cuda, name):
    """The CFL and Peclet monitor of CorUVW.f90:124-147 on the GPU, per
    material, bit for bit the oracle's (a maximum is exact in any order), on
    every mesh of the momentum tests, with fF different from the geometric f
    the monitor uses; from device and from host arrays."""
    import ctypes
    import torch
    if name == "channel_64x64x40":
        case = pkg.meshgen.calcps_case((64, 64, 40), seed=81)
        uf = pkg.meshgen.uvw_fields(case, seed=82)
    else:
        case, uf = _case(pkg, name)
    nmat = uf["nmaterials"]
    flux = np.asarray(uf["flux_u"], float) * np.linspace(-1.5, 2.0, case.ns)
    dc = pkg.krylov.DeviceCalcPS(case)
    dc.uvw_setup(uf)
    for dt, visc in ((0.01, 0.02), (0.5, 0.0)):
        want = oracle.cfl_peclet(case, uf["material"], nmat, uf["scoef"], case.rho_n, flux, dt, visc)
        rho = torch.from_numpy(np.ascontiguousarray(case.rho_n, np.float64)).cuda()
        fl = torch.from_numpy(flux).cuda()
        got = dc.cfl_peclet(rho, fl, dt, visc)
        print(f"case {name} dt {dt} visc {visc} cfl {got[0]} pe {got[1]}", flush=True)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        cfl, pe = np.zeros(nmat), np.zeros(nmat)
        rc = dc.lib.tflows_calcps_cfl_peclet(dc.handle, dt, visc, np.ascontiguousarray(case.rho_n).ctypes.data,
                                             flux.ctypes.data, cfl.ctypes.data, pe.ctypes.data, 0, None)
        assert rc == 0
        assert np.array_equal(cfl, want[0]) and np.array_equal(pe, want[1])
    assert np.all(want[0] >= 0) and np.any(want[0] > 0)
    dc.close()


def test_cfl_peclet_needs_setup(pkg, cuda):
    """Without tflows_calcps_uvw_setup there are no Scoef and materials: the
    call is refused with an error, not computed on missing arrays."""
    import torch
    case, uf = _case(pkg, "channel")
    dc = pkg.krylov.DeviceCalcPS(case)
    rho = torch.from_numpy(np.ascontiguousarray(case.rho_n, np.float64)).cuda()
    fl = torch.from_numpy(np.asarray(uf["flux_u"], float)).cuda()
    with pytest.raises(pkg.krylov.TFlowsError, match="uvw_setup"):
        dc.cfl_peclet(rho, fl, 0.1, 0.01)
    dc.close()


@pytest.mark.parametrize("shape,order", [((64, 64, 40), 2), ((64, 64, 41), 1), ((32, 64, 16), 1),
                                         ((128, 2, 4), 2), ((12, 10, 8), 0)])
def test_visit_order_selection(pkg, oracle, cuda, shape, order):
    """Which visiting order a mesh takes: blocks of 64 x 2 x 2 cells where
    lines are whole wavefronts and lines and planes come in pairs; else tiles
    brick by brick where planes are whole tiles; else tiles in order.  The
    outputs are the oracle's bit for bit in each case."""
    mg = pkg.meshgen
    case = mg.calcps_case(shape, seed=91)
    uf = mg.uvw_fields(case, seed=92)
    p = _params(pkg, "implicit", uf["nmaterials"], 2)
    dc, args, got = _run(pkg, case, uf, p)
    print(f"shape {shape} visit order {dc.visit_order()}", flush=True)
    assert dc.visit_order() == order
    want = oracle.newuvw(case, uf, p)
    for k in ("aval", "b", "rhoui"):
        assert np.array_equal(got[k], want[k]), k
    dc.close()
