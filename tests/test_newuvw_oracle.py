"""CPU pins of the momentum-assembly oracle (oracle/calcps_oracle.c,
oracle_newuvw, after Process/NewUVW.f90:137-499).

The reference routine needs the whole solver's modules (all_mod, pro_mod,
Moin_Problem_Mod, CalMinMax, bicg) and cannot be compiled on its own, and the
reference holds no vectors for it, so the restatement is pinned by:
  * hand values on a two-cell mesh with every boundary type (convection with
    the upwind part and relaxation; diffusion with cross diffusion);
  * identities the discretisation fixes:
      - convection: with the implicit upwind matrix and the deferred
        correction URFC (C - X), b - A rhoU^n is the central convective flux
        sum of every cell whatever the upwind part is;
      - diffusion: for a linear velocity field with its exact gradients,
        b - A rhoU^n vanishes (the implicit and explicit normal parts cancel
        and the stress flux of a constant gradient sums to zero over a closed
        cell) -- for U, V and W, so the argument order of
        Process/Processor.f90:314-340 is pinned too;
      - the stress flux against the tensor form
        tau = G + G^T - 2/3 tr(G) I on a single interior face;
      - inertia and relaxation: the diagonal-only system solves to the
        Euler / BDF2 step blended with the old value by URF;
  * ConvScheme's face schemes (Process/ConvScheme.f90) against their
    textbook forms on a uniform mesh, the direction taken from the mass
    flux, CDS = central on any mesh, and CalMinMax against numpy;
  * the BLEND values it refuses.
"""
import dataclasses

import numpy as np
import pytest

from test_calcps_oracle import _two_cell_case

YES, NO, FI, CN, LIN, PAR = 10, 11, 3, 2, 4, 5


def _hand_case(pkg):
    """The two-cell mesh of the pressure tests (faces: 1-2 interior, OUTFLOW
    (1,-1), INFLOW (2,-2), PRESSURE (2,-3), SYMMETRY (1,-4), BUFFER (1,-5))
    with a density on every boundary cell."""
    case = _two_cell_case(pkg)
    at = lambda c: c + case.nbc                                         # noqa: E731
    rho = np.zeros(case.nbc + case.nc + 1)
    rho[at(1)], rho[at(2)] = 1.0, 2.0
    for c in range(1, case.nbc + 1):
        rho[at(-c)] = 1.0
    u = np.zeros_like(rho)
    u[at(1)], u[at(2)], u[at(-2)], u[at(-1)] = 2.0, 1.0, 1.5, 0.75
    return dataclasses.replace(case, rho_n=rho, u=u)


def _hand_fields(case, flux, grads=None):
    n = case.nbc + case.nc + 1
    at = lambda c: c + case.nbc                                         # noqa: E731
    sx, sy, sz, dx, dy, dz = case.sx, case.sy, case.sz, case.dx, case.dy, case.dz
    with np.errstate(invalid="ignore", divide="ignore"):
        scoef = ((sx * sx + sy * sy) + sz * sz) / ((dx * sx + dy * sy) + dz * sz)
    uf = {"scoef": scoef,
          "ff": np.array([0.5, 1.0, 1.0, 1.0, 1.0, 0.5]), "material": np.ones(n, np.int32), "nmaterials": 1,
          "flux_u": np.asarray(flux, float)}
    for k in ("ux", "uy", "uz", "vx", "vy", "vz", "wx", "wy", "wz"):
        uf[k] = np.full(n, (grads or {}).get(k, 0.0))
    for k in "uvw":
        uf[k + "_o"], uf[k + "_oo"], uf["src_" + k] = np.zeros(n), np.zeros(n), np.zeros(n)
    uf["u_o"][at(1)], uf["u_o"][at(2)] = 1.0, -1.0
    uf["src_u"][at(1)] = 0.5
    return uf


def test_hand_values_convection(pkg, oracle):
    """visc = 0, BLEND = BLEND_TEM = YES, URFC = 1, LIN inertia (dt 0.5,
    V = 1), Px = (1, 3), URF = 0.5.  rhoU: cell 1 2, cell 2 2, cell -2 1.5.
    Faces (Flux_u; rhoUiS): 1-2 (3; 2), OUTFLOW (1; 2), INFLOW (-2; 1.75),
    PRESSURE (0.5; 2), SYMMETRY (0; 2), BUFFER (-1; 1).
    C = (-6 - 2 - 0 + 1, 6 + 3.5 - 1) = (-7, 8.5); upwind X = (-6 - 2, 6 + 3
    - 1) = (-8, 8); b = C - X = (1, 0.5).  Matrix: face 1-2 A21 = 3 -> diag1
    3, A(2,1) = -3, A(1,2) = 0; OUTFLOW diag1 += 1; INFLOW A12 = 2 -> b2 +=
    2 * 1.5; BUFFER A12 = 1 -> Abou(-5) = -1.  Inertia: diag += 2, b1 += 2 +
    0.5, b2 += -2; pressure b -= (1, 3): b = (2.5, -1.5), Asave = (6, 2).
    Relaxation: b += diag * 0.5 * 2 / 0.5 = (12, 4), diag / 0.5."""
    case = _hand_case(pkg)
    uf = _hand_fields(case, [3.0, 1.0, -2.0, 0.5, 0.0, -1.0])
    px = np.zeros(case.nbc + case.nc + 1)
    px[case.nbc + 1], px[case.nbc + 2] = 1.0, 3.0
    case = dataclasses.replace(case, px=px, dt=0.5)
    p = pkg.meshgen.uvw_params(visc=0.0, dt=0.5, urf=0.5, blend=[YES], blend_tem=[YES])
    out = oracle.newuvw(case, uf, p)
    # row 1: [diag, (1,2)], row 2: [(2,1), diag], trailing slot
    assert out["aval"].tolist() == [12.0, 0.0, -3.0, 4.0, 0.0]
    assert out["b"].tolist() == [14.5, 2.5]
    assert out["abou"].tolist() == [-1.0, 0.0, 0.0, 0.0, 0.0]
    assert out["asave"][case.nbc + 1:].tolist() == [6.0, 2.0]
    assert out["rhoui"].tolist() == [0.0, 0.0, 0.0, 1.5, 0.75, 0.0, 2.0, 2.0]


def test_hand_values_diffusion(pkg, oracle):
    """No flux, visc = 0.5, BLEND = NO, no inertia, URF 1 (rhoU(-1) = 0.75
    at the OUTFLOW face, which takes no diffusion term, :396-399); constant
    gradients ux = 1, uy = 2, vx = 3, wz = -1, so for U the stress flux is
    Fex = visc (4/3 + 2/3) Sx + visc (2 + 3) Sy = Sx + 2.5 Sy and
    Fim = visc Scoef (Dx + 2 Dy).  Scoef = (2, 2, 1, 2, 1, 1).
    X: face 1-2 2 - 1 = 1 (cell 1), -1 (cell 2); OUTFLOW 10 - 4; INFLOW 0;
    PRESSURE 1 - 0.5; SYMMETRY 0; BUFFER -1 + 0.5: X = (6.5, -0.5).
    D21/D12: face 1-2 A0 = 1: diag1 1, A(2,1) -1, diag2 0.5, A(1,2) -0.5;
    INFLOW A0 = 0.5: diag2 += 0.25, b2 += 0.5 * 1.5; BUFFER: diag1 += 0.5,
    Abou(-5) = -0.5.  b = (6.5, 0.75 - 0.5) - Px V = (6.5 - 1, 0.25 - 3)."""
    case = _hand_case(pkg)
    uf = _hand_fields(case, np.zeros(6), {"ux": 1.0, "uy": 2.0, "vx": 3.0, "wz": -1.0})
    p = pkg.meshgen.uvw_params(visc=0.5, inert=0, urf=1.0)
    out = oracle.newuvw(case, uf, p)
    np.testing.assert_allclose(out["aval"], [1.5, -0.5, -1.0, 0.75, 0.0], rtol=1e-15, atol=0)
    np.testing.assert_allclose(out["b"], [5.5, -2.75], rtol=1e-14, atol=0)
    assert out["abou"].tolist() == [-0.5, 0.0, 0.0, 0.0, 0.0]


def _apply(case, out, x):
    """A x with the assembled values (rows 1..NC), plus Abou(c2) x(c2) for
    BUFFER faces; x on (-NbC:NC)."""
    m = case.matrix(out["aval"]).to_scipy()
    y = m @ x[case.nbc + 1:]
    c1, c2 = case.sidec[1], case.sidec[2]
    buf = (c2 < 0) & (case.typebc[case.nbc + np.minimum(c2, 0)] == 11)
    np.add.at(y, c1[buf] - 1, out["abou"][case.nbc + c2[buf]] * x[case.nbc + c2[buf]])
    return y


CLOSED_BC = {"xmin": "inflow", "xmax": "convect", "ymin": "wall", "ymax": "wallfl", "zmin": "outflow",
             "zmax": "buffer"}
# every boundary face with implicit diffusion (OUTFLOW has none, :396-399)
DIFFUSIVE_BC = dict(CLOSED_BC, zmin="wall")


@pytest.mark.parametrize("nmat", [1, 2])
def test_convection_identity(pkg, oracle, nmat):
    """With the upwind matrix (BLEND = YES) and its explicit counterpart in b
    (BLEND_TEM = YES), URFC (C - X): b - A rhoU^n = URFC C + (1 - URFC) X per
    cell, C the central flux sum -sum_s sign Flux_u rhoUiS and X the upwind
    one, both summed here with numpy over the face list."""
    mg = pkg.meshgen
    case = mg.calcps_case((7, 6, 5), seed=41, bc=CLOSED_BC, face_seed=3)
    case = dataclasses.replace(case, px=np.zeros_like(case.px))
    uf = mg.uvw_fields(case, seed=42, nmaterials=nmat)
    urfc = [1.0, 0.625][:nmat]
    p = mg.uvw_params(nmaterials=nmat, visc=0.0, inert=0, urf=1.0, blend=[YES] * nmat, blend_tem=[YES] * nmat,
                      urfc=urfc)
    out = oracle.newuvw(case, uf, p)
    nb = case.nbc
    ru = case.rho_n * case.u
    c1, c2 = case.sidec[1].astype(int), case.sidec[2].astype(int)
    F = uf["flux_u"]
    cen = F * (case.f * ru[nb + c1] + (1 - case.f) * ru[nb + c2])
    upw = F * np.where(F < 0, ru[nb + c2], ru[nb + c1])
    Cs, Xs = np.zeros(case.nc), np.zeros(case.nc)
    inner = c2 > 0
    np.add.at(Cs, c1 - 1, -cen)
    np.add.at(Cs, c2[inner] - 1, cen[inner])
    np.add.at(Xs, c1 - 1, -upw)
    np.add.at(Xs, c2[inner] - 1, upw[inner])
    w = np.asarray(urfc)[uf["material"][nb + 1:] - 1]
    lhs = out["b"] - _apply(case, out, ru)
    scale = np.abs(cen).max() * 8
    np.testing.assert_allclose(lhs, w * Cs + (1 - w) * Xs, rtol=0, atol=1e-13 * scale)
    assert np.abs(Cs).max() > 1e-2 * scale / 8       # the identity is not vacuous


@pytest.mark.parametrize("var", [1, 2, 3])
def test_diffusion_identity_linear_field(pkg, oracle, var):
    """Linear velocity u_i = G_i . x with its exact (constant) gradients,
    varying density, no flux: the implicit diffusion (D12 = A0/rho(c2),
    D21 = A0/rho(c1)) applied to rhoU gives A0 (u1 - u2) = -Fim on each face,
    which cancels the -Fim in the cross-diffusion part, and the stress flux
    of a constant gradient sums to zero over a closed cell: b - A rhoU^n = 0
    to rounding.  A wrong pairing of Si..Dk with the gradients (var), a
    swapped D12/D21, or a sign in X breaks it."""
    mg = pkg.meshgen
    case = mg.calcps_case((6, 7, 5), seed=43, bc=DIFFUSIVE_BC, face_seed=5)
    case = dataclasses.replace(case, px=np.zeros_like(case.px), py=np.zeros_like(case.py),
                               pz=np.zeros_like(case.pz))
    G = np.random.default_rng(44).uniform(-1, 1, (3, 3))
    uf = mg.uvw_fields(case, seed=45, linear=G)
    uf["flux_u"] = np.zeros(case.ns)
    p = mg.uvw_params(var=var, visc=0.3, inert=0, urf=1.0)
    out = oracle.newuvw(case, uf, p)
    comp = "uvw"[var - 1]
    ru = case.rho_n * uf[comp]
    lhs = out["b"] - _apply(case, out, ru)
    scale = np.abs(out["aval"]).max() * np.abs(uf[comp]).max()
    assert np.abs(lhs).max() <= 1e-12 * scale
    # the stress part is there: without the cross diffusion the residual is
    # the stress flux, which a constant gradient cancels only over the
    # whole cell (not face by face)
    p0 = dict(p, cross=0)
    out0 = oracle.newuvw(case, uf, p0)
    assert np.abs(out0["b"] - _apply(case, out0, ru)).max() > 1e-6 * scale


@pytest.mark.parametrize("var", [1, 2, 3])
def test_stress_flux_tensor_form(pkg, oracle, var):
    """On the two-cell mesh with only the interior face carrying a stress
    flux (other faces have zero S and D), cell 1's b = Fex - Fim with
    Fex = visc (tau S)_i, tau = G + G^T - 2/3 tr(G) I, Fim = visc Scoef
    (G_i . D) -- the tensor written out with numpy, independent of the
    argument permutation."""
    mg = pkg.meshgen
    case = _hand_case(pkg)
    S = np.zeros((6, 3))
    D = np.zeros((6, 3))
    S[0] = [0.7, -0.4, 1.3]
    D[0] = [1.1, 0.2, -0.3]
    for k in range(1, 6):                     # boundary faces: no area, no distance
        D[k] = [1.0, 0.0, 0.0]
    case = dataclasses.replace(case, sx=S[:, 0].copy(), sy=S[:, 1].copy(), sz=S[:, 2].copy(),
                               dx=D[:, 0].copy(), dy=D[:, 1].copy(), dz=D[:, 2].copy(), px=np.zeros_like(case.px))
    G = np.array([[0.3, -1.2, 0.5], [0.8, 0.1, -0.7], [1.5, 0.4, -0.9]])
    grads = {a + b: G[i, j] for i, a in enumerate("uvw") for j, b in enumerate("xyz")}
    uf = _hand_fields(case, np.zeros(6), grads)
    uf["scoef"][1:] = 0.0                     # no area on the boundary faces
    visc = 0.37
    p = mg.uvw_params(var=var, visc=visc, inert=0, urf=1.0, diffus=0)
    out = oracle.newuvw(case, uf, p)
    i = var - 1
    tau = G + G.T - 2.0 / 3.0 * np.trace(G) * np.eye(3)
    fex = visc * tau[i] @ S[0]
    scoef = (S[0] @ S[0]) / (S[0] @ D[0])
    fim = visc * scoef * (G[i] @ D[0])
    np.testing.assert_allclose(out["b"], [fex - fim, -(fex - fim)], rtol=1e-13)


@pytest.mark.parametrize("inert", [LIN, PAR])
def test_inertia_relaxation_closed_form(pkg, oracle, inert):
    """No flux and no viscosity: the system is diagonal and solves to
    x = URF x* + (1 - URF) rhoU^n, x* the Euler step rho_o u_o + (s - Pi) dt
    (LIN) or the BDF2 step (2 rho_o u_o - rho_oo u_oo / 2 + (s - Pi) dt) / 1.5
    (PAR); Asave is V/dt or 1.5 V/dt."""
    mg = pkg.meshgen
    case = mg.calcps_case((5, 6, 4), seed=46, bc=CLOSED_BC)
    uf = mg.uvw_fields(case, seed=47)
    uf["flux_u"] = np.zeros(case.ns)
    dt, urf = 0.02, 0.7
    p = mg.uvw_params(visc=0.0, inert=inert, urf=urf, dt=dt)
    out = oracle.newuvw(case, uf, p)
    nb = case.nbc
    cells = slice(nb + 1, None)
    diag = out["aval"][case.adia - 1]
    off = np.delete(out["aval"], case.adia - 1)
    assert np.all(off == 0.0) and np.all(out["abou"] == 0.0)
    x = out["b"] / diag
    ro, uo = case.rho_o[cells], uf["u_o"][cells]
    net = (uf["src_u"][cells] - case.px[cells]) * dt
    if inert == LIN:
        xs = ro * uo + net
        V = case.volume[cells] / dt
    else:
        xs = (2.0 * ro * uo - 0.5 * case.rho_oo[cells] * uf["u_oo"][cells] + net) / 1.5
        V = 1.5 * case.volume[cells] / dt
    want = urf * xs + (1 - urf) * (case.rho_n * case.u)[cells]
    np.testing.assert_allclose(x, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    np.testing.assert_allclose(out["asave"][cells], V, rtol=1e-15)


CDS, QUICK, LUDS, MINMOD, SMART, AVL_SMART, SUPERBEE, GAMMA = range(20, 28)   # Process/IniPar.f90:42-49
SCHEMES = [CDS, QUICK, LUDS, MINMOD, SMART, AVL_SMART, SUPERBEE, GAMMA]


def one_face_case(pkg, triple, flux_sign, mass_sign=None):
    """A uniform 8 x 3 x 3 box (f = 1/2 on interior faces, spacing h in x)
    with Flux_u nonzero (flux_sign) on one interior x-face only, between the
    cells at x index 3 (c1) and 4 (c2), and a field phi of x index alone:
    (phi_U, phi_C, phi_D) = triple on (2, 3, 4) when the mass flux (of sign
    mass_sign, default flux_sign) runs from c1 to c2, on (5, 4, 3) when it
    runs back; other cells take phi_C.  The x gradient of every cell is the
    central difference of its x neighbours, so ConvScheme's extrapolated
    upstream value phi_D -/+ 2 grad . d lands on phi_U.  Returns (case,
    face, c1, c2, phi on (-NbC:NC), d phi/dx)."""
    mg = pkg.meshgen
    case = mg.calcps_case((8, 3, 3), seed=5, ratio=1.0, perturb=0.0)
    nb = case.nbc
    mass_sign = flux_sign if mass_sign is None else mass_sign
    h = 2.0 * np.pi / 8
    ix = np.clip(np.floor(case.xc[:, 0] / h).astype(int), 0, 7)
    table = np.full(8, float(triple[1]))
    idx = (2, 3, 4) if mass_sign > 0 else (5, 4, 3)
    for i, v in zip(idx, triple):
        table[i] = v
    phi = table[ix]
    grad = (table[np.minimum(ix + 1, 7)] - table[np.maximum(ix - 1, 0)]) / (2.0 * h)
    c1s, c2s = case.sidec[1], case.sidec[2]
    inner = c2s > 0
    mid = lambda c: (ix[nb + c] , np.rint(case.xc[nb + c, 1] * 1e6), np.rint(case.xc[nb + c, 2] * 1e6))  # noqa: E731
    cand = [s for s in np.nonzero(inner & (np.abs(case.sx) > 0))[0]
            if ix[nb + c1s[s]] == 3 and ix[nb + c2s[s]] == 4]
    face = int(cand[len(cand) // 2])
    return case, face, int(c1s[face]), int(c2s[face]), phi, grad


def _psi(scheme, r):
    """The limiters in the textbook TVD form phi_f = phi_C + psi(r)/2
    (phi_D - phi_C), r = (phi_C - phi_U) / (phi_D - phi_C), on a uniform
    mesh: QUICK 3/4 + r/4, linear upwind r, minmod, superbee (Roe), SMART
    (Gaskell & Lau) min(4r, 3/4 + r/4, 2), central 1."""
    return {CDS: 1.0, QUICK: 0.75 + 0.25 * r, LUDS: r, MINMOD: max(0.0, min(r, 1.0)),
            SUPERBEE: max(0.0, min(2 * r, 1.0), min(r, 2.0)), SMART: max(0.0, min(4 * r, 0.75 + 0.25 * r, 2.0)),
            YES: 1.0}[scheme]


def expected_face_value(scheme, triple):
    """phi_f of a scheme for (phi_U, phi_C, phi_D) on a uniform mesh, from
    the textbook forms: _psi; AVL-SMART equal to SMART wherever SMART is not
    bounded by 2 (the two share the QUICK branch), upwind for r <= 0, and
    bounded by fj (2 - fj) = 3/4 of the step above; the Gamma scheme (Jasak)
    on the normalised phi~_C = (phi_C - phi_U) / (phi_D - phi_U): upwind
    outside (0, 1), central on [beta, 1), the blend (1 - g/2) phi_C +
    g/2 phi_D with g = phi~_C / beta below beta = 0.1."""
    pu, pc, pd = triple
    r = (pc - pu) / (pd - pc)
    if scheme == GAMMA:
        nc = (pc - pu) / (pd - pu)
        if nc <= 0.0 or nc >= 1.0:
            return pc
        if nc >= 0.1:
            return 0.5 * (pc + pd)
        g = nc / 0.1
        return (1 - g / 2) * pc + g / 2 * pd
    if scheme == AVL_SMART:
        half = 0.0 if r <= 0 else min(1.25 * r, 0.375 + 0.125 * r, 0.75)
        return pc + half * (pd - pc)
    return pc + 0.5 * _psi(scheme, r) * (pd - pc)


# (phi_U, phi_C, phi_D): r = -0.5 (a maximum at C), 1 (linear), 2, 10, 0.05
TRIPLES = [(-1.0, 0.0, -2.0), (1.0, 2.0, 3.0), (0.0, 2.0, 3.0), (0.0, 10.0, 11.0), (0.0, 0.05, 1.05)]


def _uvw_face_value(pkg, oracle, scheme, triple, flux_sign, mass_sign=None, grad_scale=1.0):
    mg = pkg.meshgen
    case, face, c1, c2, phi, grad = one_face_case(pkg, triple, flux_sign, mass_sign)
    grad = grad * grad_scale
    n = case.nbc + case.nc + 1
    case = dataclasses.replace(case, rho_n=np.ones(n), px=np.zeros(n))
    uf = mg.uvw_fields(case, seed=9)
    uf["u"] = phi
    uf["ux"], uf["uy"], uf["uz"] = grad, np.zeros(n), np.zeros(n)
    uf["u_o"], uf["u_oo"], uf["src_u"] = np.zeros(n), np.zeros(n), np.zeros(n)
    F = 0.75 * flux_sign
    uf["flux_u"] = np.zeros(case.ns)
    uf["flux_u"][face] = F
    if mass_sign is not None:
        uf["flux"] = np.zeros(case.ns)
        uf["flux"][face] = 0.5 * mass_sign
    p = mg.uvw_params(var=1, visc=0.0, inert=0, urf=1.0, blend=[scheme], blend_tem=[NO], urfc=[1.0])
    out = oracle.newuvw(case, uf, p)
    # only this face convects: b(c1) = -F phi_f, b(c2) = +F phi_f
    assert out["b"][c1 - 1] == pytest.approx(-out["b"][c2 - 1], rel=1e-15)
    return out["b"][c2 - 1] / F


@pytest.mark.parametrize("scheme", SCHEMES + [YES])
@pytest.mark.parametrize("flux_sign", [1.0, -1.0])
def test_conv_scheme_textbook_face_values(pkg, oracle, scheme, flux_sign):
    """ConvScheme (Process/ConvScheme.f90:55-135) through the assembly: the
    face value of each scheme, in both flow directions, for an extremum at
    the upwind cell, a linear profile, and steep and shallow ones, against
    the textbook forms of the schemes on a uniform mesh."""
    for triple in TRIPLES:
        got = _uvw_face_value(pkg, oracle, scheme, triple, flux_sign)
        want = expected_face_value(scheme, triple) if scheme != YES else 0.5 * (triple[1] + triple[2])
        assert got == pytest.approx(want, rel=1e-12, abs=1e-13), (scheme, triple)


@pytest.mark.parametrize("scheme", [LUDS, SUPERBEE, GAMMA])
def test_conv_scheme_direction_is_the_mass_flux(pkg, oracle, scheme):
    """The upwind cell follows the module's mass flux Flux (ConvScheme.f90:
    57), not Flux_u, which only weighs the face value into C: with the two of
    opposite sign the face value is the one of the mass flux's direction."""
    for triple in TRIPLES[2:]:
        for fs in (1.0, -1.0):
            got = _uvw_face_value(pkg, oracle, scheme, triple, fs, mass_sign=-fs)
            assert got == pytest.approx(expected_face_value(scheme, triple), rel=1e-12), (scheme, triple, fs)


@pytest.mark.parametrize("scheme", [QUICK, LUDS, SMART, SUPERBEE])
def test_conv_scheme_upstream_value_is_bounded(pkg, oracle, scheme):
    """The extrapolated upstream value is limited to the upwind cell's
    PHImin .. PHImax (ConvScheme.f90:77): with the upwind cell's gradient
    ten times too steep it overshoots far beyond its neighbours and is
    brought back to the nearest bound -- here phi_U's own value, so the
    face value is the textbook one again.  A ten times too flat gradient
    stays inside the bounds and is not limited, so the value changes."""
    for triple in ((0.0, 2.0, 3.0), (1.0, 2.0, 3.0), (0.0, 10.0, 11.0)):
        for fs in (1.0, -1.0):
            got = _uvw_face_value(pkg, oracle, scheme, triple, fs, grad_scale=10.0)
            assert got == pytest.approx(expected_face_value(scheme, triple), rel=1e-12), (scheme, triple, fs)
            flat = _uvw_face_value(pkg, oracle, scheme, triple, fs, grad_scale=0.1)
            assert abs(flat - expected_face_value(scheme, triple)) > 1e-6


@pytest.mark.parametrize("blend", [[CDS], [CDS, CDS], [GAMMA, YES]])
def test_cds_keeps_the_convection_identity(pkg, oracle, blend):
    """CDS gives the central value f phi1 + (1 - f) phi2 in both directions
    on any mesh, so on the stretched, perturbed box (f != 1/2) the assembly
    with CDS equals the one with YES (BLEND /= NO on every face either way,
    so the implicit upwind part is the same).  With GAMMA it does not."""
    mg = pkg.meshgen
    nmat = len(blend)
    case = mg.calcps_case((7, 6, 5), seed=41, bc=CLOSED_BC, face_seed=3)
    case = dataclasses.replace(case, px=np.zeros_like(case.px))
    uf = mg.uvw_fields(case, seed=42, nmaterials=nmat)
    p = mg.uvw_params(nmaterials=nmat, visc=0.0, inert=0, urf=1.0, blend=blend, blend_tem=[YES] * nmat,
                      urfc=[1.0] * nmat)
    ref = oracle.newuvw(case, uf, dict(p, blend=[YES] * nmat))
    out = oracle.newuvw(case, uf, p)
    scale = np.abs(ref["b"]).max()
    if blend[0] == CDS:
        np.testing.assert_allclose(out["b"], ref["b"], rtol=0, atol=1e-13 * scale)
        assert np.array_equal(out["aval"], ref["aval"])
    else:
        assert np.abs(out["b"] - ref["b"]).max() > 1e-6 * scale


def test_calminmax_brute_force(pkg, oracle):
    """Process/CalMinMax.f90: each cell's max / min over itself and the
    cells across its interior and BUFFER faces, against numpy's
    maximum.at / minimum.at over the face list."""
    mg = pkg.meshgen
    case = mg.calcps_case((6, 5, 4), seed=46, bc=CLOSED_BC, face_seed=4)
    phi = np.random.default_rng(47).standard_normal(case.nbc + case.nc + 1)
    mx, mn = oracle.calminmax(case, phi)
    nb = case.nbc
    c1, c2 = case.sidec[1].astype(int), case.sidec[2].astype(int)
    use = (c2 > 0) | ((c2 < 0) & (case.typebc[nb + np.minimum(c2, 0)] == 11))
    wmx, wmn = phi.copy(), phi.copy()
    for a, b in ((c1, c2), (c2, c1)):
        np.maximum.at(wmx, nb + a[use], phi[nb + b[use]])
        np.minimum.at(wmn, nb + a[use], phi[nb + b[use]])
    assert np.array_equal(mx, wmx) and np.array_equal(mn, wmn)
    assert np.any(use & (c2 < 0)) and np.any(mx != phi)


def test_blend_schemes_refused(pkg, oracle):
    """YES next to NO at a face (where the reference's ConvScheme leaves
    Phij undefined) and codes that are no scheme are refused; YES on both
    materials and every scheme are not."""
    mg = pkg.meshgen
    case = mg.calcps_case((6, 4, 3), seed=48)
    uf = mg.uvw_fields(case, seed=49, nmaterials=2)
    for bad in ([YES, NO], [28, 28], [19, 19]):
        with pytest.raises(ValueError):
            oracle.newuvw(case, uf, mg.uvw_params(nmaterials=2, blend=bad))
    oracle.newuvw(case, uf, mg.uvw_params(nmaterials=2, blend=[YES, YES]))
    for sch in SCHEMES:
        oracle.newuvw(case, uf, mg.uvw_params(nmaterials=2, blend=[sch, NO]))


def _rt_case(pkg):
    """The hand mesh with densities and fF such that every face has RhoS =
    fF rho(c1) + (1 - fF) rho(c2) = 2.5 exactly: rho = 1 and 3 on the cells,
    4 on the boundary cells of cell 1 (-1 OUTFLOW, -4 SYMMETRY, -5 BUFFER), 2
    on those of cell 2 (-2 INFLOW, -3 PRESSURE); fF 0.25 at the interior face
    and 0.5 at the others.  With the sides swapped the interior face would
    give 1.5, with the geometric f instead of fF something else again."""
    case = _hand_case(pkg)
    at = lambda c: c + case.nbc                                         # noqa: E731
    rho = case.rho_n.copy()
    rho[at(1)], rho[at(2)] = 1.0, 3.0
    for c, r in ((-1, 4.0), (-2, 2.0), (-3, 2.0), (-4, 4.0), (-5, 4.0)):
        rho[at(c)] = r
    case = dataclasses.replace(case, rho_n=rho, dt=0.5)
    grads = {"ux": 1.0, "uy": -2.0, "uz": 0.5, "vx": 3.0, "vy": 0.25, "wx": -1.0, "wz": 2.0}
    uf = _hand_fields(case, [3.0, 1.0, -2.0, 0.5, 0.0, -1.0], grads)
    uf["ff"] = np.array([0.25, 0.5, 0.5, 0.5, 0.5, 0.5])
    assert not np.array_equal(uf["ff"], case.f)
    return case, uf


@pytest.mark.parametrize("var", [1, 2, 3])
def test_density_weighted_viscosity(pkg, oracle, var):
    """The Moin_Problem() == 4 branch (NewUVW.f90:285-295): VISeff = VISc
    RhoS.  Where every face has RhoS = 2.5, the assembly with visc_rho = 1
    and VISc = 0.125 is bit for bit the plain one with VISc = 0.3125 (both
    multiply to the same double): the branch scales both Fex and A0, with fF
    and the sides in the reference's roles.  Without the branch the
    diffusion terms differ."""
    mg = pkg.meshgen
    case, uf = _rt_case(pkg)
    base = dict(var=var, visc=0.125, dt=0.5, urf=0.5, blend=[YES], blend_tem=[YES])
    rt = oracle.newuvw(case, uf, mg.uvw_params(visc_rho=1, **base))
    scaled = oracle.newuvw(case, uf, mg.uvw_params(**dict(base, visc=0.3125)))
    plain = oracle.newuvw(case, uf, mg.uvw_params(**base))
    for k in ("aval", "b", "abou", "asave", "rhoui"):
        assert np.array_equal(rt[k], scaled[k]), k
    assert not np.array_equal(rt["aval"], plain["aval"])
    assert not np.array_equal(rt["b"], plain["b"])

