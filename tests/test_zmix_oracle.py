"""CPU pins of the mixture-fraction assembly oracle (oracle/calcps_oracle.c,
oracle_calczmix / oracle_zmix_update, after Process/CalcZmix.f90:100-405).

Like NewUVW, the reference routine needs the solver's modules and cannot be
built alone, and the reference holds no vectors for it.  Pins: hand values
on the two-cell mesh (every boundary type), the order of the two-level
inertia product (V/dt Z_o rho_o, which differs from the momentum equation's
V/dt rho_o U_o), the convection identity with the upwind matrix, the
diffusion identity for a linear field, the inertia / relaxation closed form,
the [0, 1] limiter after the solve, ConvScheme's face schemes against their
textbook forms (both directions, the mass flux's), and the BLEND_TEM values
it refuses.
"""
import dataclasses

import numpy as np
import pytest

from test_newuvw_oracle import (CLOSED_BC, DIFFUSIVE_BC, GAMMA, SCHEMES, TRIPLES, _apply, _hand_case, _hand_fields,
                                expected_face_value, one_face_case)

YES, NO, LIN, PAR = 10, 11, 4, 5


def _hand_zfields(case, flux):
    uf = _hand_fields(case, flux)
    zf = {k: uf[k] for k in ("scoef", "ff", "material", "nmaterials", "flux_u")}
    zf["z"] = case.u.copy()                       # z: cell 1 2, cell 2 1, cell -2 1.5, cell -1 0.75
    zf["z_o"], zf["z_oo"], zf["src_z"] = uf["u_o"], uf["u_oo"], uf["src_u"]
    for k in ("zx", "zy", "zz"):
        zf[k] = np.zeros_like(zf["z"])
    return zf


def test_hand_values(pkg, oracle):
    """The convection hand case of the momentum equation (BLEND_TEM = YES
    now drives both the matrix and the upwind part), without the pressure
    term and Asave: before relaxation b = (3.5, 1.5), diagonal (6, 2); URF
    0.5 adds diag rhoZ = (12, 4)."""
    case = dataclasses.replace(_hand_case(pkg), dt=0.5)
    zf = _hand_zfields(case, [3.0, 1.0, -2.0, 0.5, 0.0, -1.0])
    p = pkg.meshgen.uvw_params(visc=0.0, dt=0.5, urf=0.5, blend_tem=[YES])
    out = oracle.calczmix(case, zf, p)
    assert out["aval"].tolist() == [12.0, 0.0, -3.0, 4.0, 0.0]
    assert out["b"].tolist() == [15.5, 5.5]
    assert out["abou"].tolist() == [-1.0, 0.0, 0.0, 0.0, 0.0]
    assert out["rhoz"].tolist() == [0.0, 0.0, 0.0, 1.5, 0.75, 0.0, 2.0, 2.0]
    # BLEND (the momentum flag) plays no part
    out2 = oracle.calczmix(case, zf, dict(p, blend=[NO]))
    assert np.array_equal(out2["b"], out["b"])


def test_two_level_inertia_order(pkg, oracle):
    """CalcZmix.f90:371 multiplies V/dt * Z_o * rho_o left to right; with
    V/dt = 0.1, Z_o = 0.11, rho_o = 0.7 that is 0.0077, where the momentum
    order V/dt * rho_o * Z_o gives 0.007699999999999999."""
    case = _hand_case(pkg)
    n = case.nbc + case.nc + 1
    rho_o = np.zeros(n)
    rho_o[case.nbc + 1] = 0.7
    case = dataclasses.replace(case, rho_o=rho_o, dt=10.0)
    zf = _hand_zfields(case, np.zeros(6))
    zf["z_o"] = np.zeros(n)
    zf["z_o"][case.nbc + 1] = 0.11
    zf["src_z"] = np.zeros(n)
    p = pkg.meshgen.uvw_params(visc=0.0, dt=10.0, urf=1.0, inert=LIN)
    b = oracle.calczmix(case, zf, p)["b"]
    assert b[0] == (0.1 * 0.11) * 0.7 == 0.0077
    assert b[0] != (0.1 * 0.7) * 0.11


@pytest.mark.parametrize("nmat", [1, 2])
def test_convection_identity(pkg, oracle, nmat):
    """b - A rhoZ^n = URFC C + (1 - URFC) X per cell (C central, X upwind
    flux sums), BLEND_TEM = YES, no diffusion or inertia."""
    mg = pkg.meshgen
    case = mg.calcps_case((7, 5, 6), seed=61, bc=CLOSED_BC, face_seed=2)
    zf = mg.zmix_fields(case, seed=62, nmaterials=nmat)
    urfc = [1.0, 0.75][:nmat]
    p = mg.uvw_params(nmaterials=nmat, visc=0.0, inert=0, urf=1.0, blend_tem=[YES] * nmat, urfc=urfc)
    out = oracle.calczmix(case, zf, p)
    nb = case.nbc
    rz = case.rho_n * zf["z"]
    c1, c2 = case.sidec[1].astype(int), case.sidec[2].astype(int)
    F = zf["flux_u"]
    cen = F * (case.f * rz[nb + c1] + (1 - case.f) * rz[nb + c2])
    upw = F * np.where(F < 0, rz[nb + c2], rz[nb + c1])
    Cs, Xs = np.zeros(case.nc), np.zeros(case.nc)
    inner = c2 > 0
    np.add.at(Cs, c1 - 1, -cen)
    np.add.at(Cs, c2[inner] - 1, cen[inner])
    np.add.at(Xs, c1 - 1, -upw)
    np.add.at(Xs, c2[inner] - 1, upw[inner])
    w = np.asarray(urfc)[zf["material"][nb + 1:] - 1]
    lhs = out["b"] - _apply(case, out, rz)
    np.testing.assert_allclose(lhs, w * Cs + (1 - w) * Xs, rtol=0, atol=1e-13 * 8 * np.abs(cen).max())


def test_diffusion_identity_linear_field(pkg, oracle):
    """z = 0.5 + g . x with its exact gradients, varying density, no flux:
    b - A rhoZ^n = 0 to rounding (the implicit part cancels -Fim face by
    face, and VISc g . S sums to zero over each closed cell); without the
    cross-diffusion part the residual is the flux g . S of the boundary
    faces, which is not zero."""
    mg = pkg.meshgen
    case = mg.calcps_case((6, 6, 7), seed=63, bc=DIFFUSIVE_BC, face_seed=4)
    zf = mg.zmix_fields(case, seed=64, linear=np.array([0.11, -0.07, 0.05]))
    zf["flux_u"] = np.zeros(case.ns)
    p = mg.uvw_params(visc=0.2, inert=0, urf=1.0)
    out = oracle.calczmix(case, zf, p)
    rz = case.rho_n * zf["z"]
    scale = np.abs(out["aval"]).max() * np.abs(zf["z"]).max()
    assert np.abs(out["b"] - _apply(case, out, rz)).max() <= 1e-12 * scale
    out0 = oracle.calczmix(case, zf, dict(p, cross=0))
    assert np.abs(out0["b"] - _apply(case, out0, rz)).max() > 1e-6 * scale


def test_flux_single_face(pkg, oracle):
    """Only the interior face has an area: cell 1's b is Fex - Fim with
    Fex = VISc g . S and Fim = VISc Scoef g . D (g the face gradient), and
    cell 2's is the negative -- each component of g against its own
    component of S and D."""
    case = _hand_case(pkg)
    S = np.zeros((6, 3))
    D = np.zeros((6, 3))
    S[0], D[0] = [0.7, -0.4, 1.3], [1.1, 0.2, -0.3]
    D[1:] = [1.0, 0.0, 0.0]
    case = dataclasses.replace(case, sx=S[:, 0].copy(), sy=S[:, 1].copy(), sz=S[:, 2].copy(), dx=D[:, 0].copy(),
                               dy=D[:, 1].copy(), dz=D[:, 2].copy())
    zf = _hand_zfields(case, np.zeros(6))
    zf["scoef"][1:] = 0.0
    g = np.array([0.3, -1.2, 0.5])
    for k, gk in zip(("zx", "zy", "zz"), g):
        zf[k] = np.full_like(zf["z"], gk)
    visc = 0.37
    out = oracle.calczmix(case, zf, pkg.meshgen.uvw_params(visc=visc, inert=0, urf=1.0, diffus=0))
    scoef = (S[0] @ S[0]) / (S[0] @ D[0])
    want = visc * (g @ S[0]) - visc * scoef * (g @ D[0])
    np.testing.assert_allclose(out["b"], [want, -want], rtol=1e-13)


@pytest.mark.parametrize("inert", [LIN, PAR])
def test_inertia_relaxation_closed_form(pkg, oracle, inert):
    """No flux, no diffusion: x = URF x* + (1 - URF) rhoZ^n with x* the
    Euler step rho_o z_o + s dt or the BDF2 step."""
    mg = pkg.meshgen
    case = mg.calcps_case((5, 4, 6), seed=65, bc=CLOSED_BC)
    zf = mg.zmix_fields(case, seed=66)
    zf["flux_u"] = np.zeros(case.ns)
    dt, urf = 0.05, 0.6
    p = mg.uvw_params(visc=0.0, inert=inert, urf=urf, dt=dt)
    out = oracle.calczmix(case, zf, p)
    nb = case.nbc
    cells = slice(nb + 1, None)
    x = out["b"] / out["aval"][case.adia - 1]
    ro, zo = case.rho_o[cells], zf["z_o"][cells]
    net = zf["src_z"][cells] * dt
    xs = ro * zo + net if inert == LIN else (2 * ro * zo - 0.5 * case.rho_oo[cells] * zf["z_oo"][cells] + net) / 1.5
    want = urf * xs + (1 - urf) * (case.rho_n * zf["z"])[cells]
    np.testing.assert_allclose(x, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


def test_update_limits_to_unit_interval(pkg, oracle):
    """Zmix%n = rhoZ / rho, then max(., 0) and min(., 1) (:401-405);
    boundary values untouched."""
    case = _hand_case(pkg)
    n = case.nbc + case.nc + 1
    rho = case.rho_n
    rhoz = np.zeros(n)
    rhoz[case.nbc + 1], rhoz[case.nbc + 2] = -0.25, 3.0          # / rho (1, 2): -0.25, 1.5
    z = np.full(n, 7.0)
    out = oracle.zmix_update(case, rhoz, rho, z)
    assert out[case.nbc + 1:].tolist() == [0.0, 1.0]
    assert np.all(out[:case.nbc + 1] == 7.0)
    rhoz[case.nbc + 1], rhoz[case.nbc + 2] = 0.5, 1.0
    assert oracle.zmix_update(case, rhoz, rho, z)[case.nbc + 1:].tolist() == [0.5, 0.5]


def _zmix_face_value(pkg, oracle, scheme, triple, flux_sign, mass_sign=None):
    mg = pkg.meshgen
    case, face, c1, c2, phi, grad = one_face_case(pkg, triple, flux_sign, mass_sign)
    n = case.nbc + case.nc + 1
    case = dataclasses.replace(case, rho_n=np.ones(n))
    zf = mg.zmix_fields(case, seed=10)
    zf["z"] = phi
    zf["zx"], zf["zy"], zf["zz"] = grad, np.zeros(n), np.zeros(n)
    zf["z_o"], zf["z_oo"], zf["src_z"] = np.zeros(n), np.zeros(n), np.zeros(n)
    F = 0.75 * flux_sign
    zf["flux_u"] = np.zeros(case.ns)
    zf["flux_u"][face] = F
    if mass_sign is not None:
        zf["flux"] = np.zeros(case.ns)
        zf["flux"][face] = 0.5 * mass_sign
    p = mg.uvw_params(visc=0.0, inert=0, urf=1.0, blend_tem=[scheme], urfc=[1.0])
    out = oracle.calczmix(case, zf, p)
    # b = C - X with the upwind part X of the same face (Flux_u's upwind cell)
    up = phi[case.nbc + (c1 if F >= 0 else c2)]
    assert out["b"][c1 - 1] == pytest.approx(-out["b"][c2 - 1], rel=1e-15)
    return out["b"][c2 - 1] / F + up


@pytest.mark.parametrize("scheme", SCHEMES + [YES])
@pytest.mark.parametrize("flux_sign", [1.0, -1.0])
def test_conv_scheme_textbook_face_values(pkg, oracle, scheme, flux_sign):
    """ConvScheme through the scalar assembly (CalcZmix.f90:163-166, with
    rhoZ and the gradients dZdx, dZdy, dZdz and Dx, Dy, Dz): each scheme's
    face value against its textbook form, both directions."""
    for triple in TRIPLES:
        got = _zmix_face_value(pkg, oracle, scheme, triple, flux_sign)
        want = expected_face_value(scheme, triple) if scheme != YES else 0.5 * (triple[1] + triple[2])
        assert got == pytest.approx(want, rel=1e-12, abs=1e-13), (scheme, triple)


def test_conv_scheme_direction_is_the_mass_flux(pkg, oracle):
    for triple in TRIPLES[2:]:
        for fs in (1.0, -1.0):
            got = _zmix_face_value(pkg, oracle, GAMMA, triple, fs, mass_sign=-fs)
            assert got == pytest.approx(expected_face_value(GAMMA, triple), rel=1e-12)


def test_blend_tem_schemes_refused(pkg, oracle):
    mg = pkg.meshgen
    case = mg.calcps_case((5, 4, 3), seed=67)
    zf = mg.zmix_fields(case, seed=68, nmaterials=2)
    for bad in ([YES, NO], [28, 28], [12, 12]):
        with pytest.raises(ValueError):
            oracle.calczmix(case, zf, mg.uvw_params(nmaterials=2, blend_tem=bad))
    oracle.calczmix(case, zf, mg.uvw_params(nmaterials=2, blend_tem=[NO, NO], blend=[21, 21]))
    for sch in SCHEMES:
        oracle.calczmix(case, zf, mg.uvw_params(nmaterials=2, blend_tem=[sch, NO]))


def test_density_weighted_viscosity(pkg, oracle):
    """The Moin_Problem() == 4 branch of CalcZmix.f90:232-242 (VISeff =
    VIScZmix RhoS) on the mesh of the momentum test where every face has
    RhoS = 2.5: visc_rho = 1 with 0.125 is bit for bit the plain assembly
    with 0.3125, and differs from the plain one with 0.125."""
    from test_newuvw_oracle import _rt_case
    mg = pkg.meshgen
    case, uf = _rt_case(pkg)
    zf = _hand_zfields(case, uf["flux_u"])
    zf["ff"] = uf["ff"]
    for k, g in zip(("zx", "zy", "zz"), (0.5, -1.0, 0.25)):
        zf[k] = np.full_like(zf["z"], g)
    base = dict(visc=0.125, dt=0.5, urf=0.5, blend_tem=[YES])
    rt = oracle.calczmix(case, zf, mg.uvw_params(visc_rho=1, **base))
    scaled = oracle.calczmix(case, zf, mg.uvw_params(**dict(base, visc=0.3125)))
    plain = oracle.calczmix(case, zf, mg.uvw_params(**base))
    for k in ("aval", "b", "abou", "rhoz"):
        assert np.array_equal(rt[k], scaled[k]), k
    assert not np.array_equal(rt["aval"], plain["aval"])
    assert not np.array_equal(rt["b"], plain["b"])

