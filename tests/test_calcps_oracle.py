"""Pins of the C restatement of the pressure-correction assembly
(oracle/calcps_oracle.c, after Process/CalcPS.f90:51-202).

The reference routine cannot be built here (its Moin_Problem_mod chain
needs gfortran's `irand`), so the restatement is pinned by:

* a two-cell mesh with one face of every kind, whose every output was worked
  out by hand in exact binary fractions (signs, which cell takes which
  term, the 0.5 and V/Asave factors, the boundary branches, Abou, the time
  term, the volume-weighted mean);
* closed forms on an orthogonal box: with V/Asave = 1 the off-diagonals are
  -|S|^2/(S.d) = -A_f/d_f, which the independently written pressure_matrix
  generator gives from the node spacings;
* invariants on a perturbed, shuffled mesh with every boundary type:
  symmetric off-diagonals, zero row sums of cells with only interior faces,
  a volume-weighted mean of zero, face-order independence, and Nmat = 2
  doubling the face-loop terms.
"""
import dataclasses

import numpy as np
import pytest


def _two_cell_case(pkg):
    """Cells 1, 2; faces: 1-2 interior, then OUTFLOW, INFLOW, PRESSURE,
    SYMMETRY and BUFFER boundary faces (boundary cells -1 .. -5)."""
    mg = pkg.meshgen
    nbc, nc = 5, 2
    topo = mg.topolm(nc, np.array([1]), np.array([2]))
    ns = 6
    sidec = np.zeros((3, ns), np.int32)
    sidec[1] = [1, 1, 2, 2, 1, 1]
    sidec[2] = [2, -1, -2, -3, -4, -5]
    sidaij = np.zeros((2, ns), np.int32)
    sidaij[:, 0] = (topo.sidaij[:, 0] + 1)
    S = np.array([[2, 0, 0], [0, 4, 0], [0, 0, -1], [1, 0, 0], [0, 0, 1], [-1, 0, 0]], float)
    D = np.array([[1, 0, 0], [0, 2, 0], [0, 0, -1], [0.5, 0, 0], [0, 0, 1], [-1, 0, 0]], float)
    f = np.array([0.5, 1.0, 0.5, 1.0, 1.0, 0.5])
    at = lambda c: c + nbc                                              # noqa: E731
    z = lambda: np.zeros(nbc + nc + 1)                                  # noqa: E731
    typebc, vol, asave = np.zeros(nbc + nc + 1, np.int32), z(), z()
    rho, rho_o, rho_oo, u, v, w, p, px, py, pz = (z() for _ in range(10))
    for c, t in zip((-1, -2, -3, -4, -5), (3, 1, 12, 4, 11)):
        typebc[at(c)] = t
    vol[at(1)], vol[at(2)], vol[at(-5)] = 1.0, 1.0, 2.0
    asave[at(1)], asave[at(2)], asave[at(-5)] = 2.0, 0.5, 1.0
    rho[at(1)], rho[at(2)], rho[at(-2)], rho[at(-5)] = 1.0, 2.0, 1.0, 1.0
    rho_o[at(1)], rho_oo[at(1)] = 1.0, 1.0
    rho_o[at(2)], rho_oo[at(2)] = 1.0, 0.0
    u[at(1)], u[at(2)] = 2.0, 1.0
    v[at(1)] = 0.5
    w[at(-2)] = -4.0
    p[at(1)], p[at(2)], p[at(-5)] = 3.0, 1.0, 3.0
    px[at(1)], px[at(2)], px[at(-5)] = 1.0, 3.0, -1.0
    return mg.CalcPSCase((2, 1, 1), nc, nbc, ns, topo.nonz, 1, 0.5, sidec, sidaij, topo.acol, topo.arow,
                         topo.adia, typebc, S[:, 0].copy(), S[:, 1].copy(), S[:, 2].copy(), D[:, 0].copy(),
                         D[:, 1].copy(), D[:, 2].copy(), f, vol, asave, rho, rho_o, rho_oo, u, v, w, p, px, py, pz)


def test_two_cell_hand_values(pkg, oracle):
    """Interior face: SMDPN = 4/2 = 2, A12 = 0.5*2*(1/2 + 1/0.5) = 2.5,
    Urs = 0.5*1*2 + 0.5*2*1 = 2, dPxi = 0.5*(1+3)*1 = 2, so
    Flux = 2*2 + 2.5*(3-1) + 2.5*2 = 14.  OUTFLOW: A12 = 2*1/2 = 1,
    Flux = Vrs*Sy = 0.5*4 = 2.  INFLOW: Wrs = 0.5*1*(-4) = -2, Flux =
    (-2)(-1) = 2, no diagonal term.  PRESSURE: A12 = 2*1/0.5 = 4, Flux =
    rho2*U2*Sx = 2.  SYMMETRY: Flux = 0.  BUFFER: SMDPN = 1,
    A12 = 0.5*(1/2 + 2/1) = 1.25, Abou(-5) = -1.25, Flux = 1*(-1) = -1.
    b before the time term: cell 1 = -14 - 2 + 1 = -15, cell 2 = 14 - 2 - 2
    = 10; errmax 15.  Time term of cell 2: V/dt*(1.5*2 - 2*1 + 0) = 2, so
    b2 = 8; bc_sum = -7, vol = 2, mean -3.5: b = (-11.5, 11.5)."""
    case = _two_cell_case(pkg)
    out = oracle.calcps(case)
    # row 1: [diag, (1,2)], row 2: [(2,1), diag], trailing slot
    assert out["aval"].tolist() == [4.75, -2.5, -2.5, 6.5, 0.0]
    assert out["flux"].tolist() == [14.0, 2.0, 2.0, 2.0, 0.0, -1.0]
    # Abou(-NbC:-1): Abou(-5) first
    assert out["abou"].tolist() == [-1.25, 0.0, 0.0, 0.0, 0.0]
    assert out["errmax"] == 15.0
    assert out["mean"] == -3.5
    assert out["b"].tolist() == [-11.5, 11.5]


def test_two_cell_nmat_twice(pkg, oracle):
    """Nmat = 2 runs the face loop and the time-term loop twice
    (CalcPS.f90:68, 188): diagonals and b accumulate twice, the
    off-diagonals, Flux and Abou are assigned, not accumulated."""
    case = _two_cell_case(pkg)
    out = oracle.calcps(case, nmat=2)
    assert out["aval"].tolist() == [9.5, -2.5, -2.5, 13.0, 0.0]
    assert out["flux"].tolist() == [14.0, 2.0, 2.0, 2.0, 0.0, -1.0]
    assert out["errmax"] == 30.0
    # b = (-30, 20) - 2 time terms (0, 2 each): (-30, 16); bc_sum is summed
    # after each pass: (-30 + 18) + (-30 + 16) = -26, vol = 4
    assert out["mean"] == -6.5
    assert out["b"].tolist() == [-23.5, 22.5]


def test_orthogonal_box_closed_form(pkg, oracle):
    """Unperturbed channel box with V/Asave = 1: A12 = |S|^2/(S.d), which
    for an orthogonal face is A_f/d_f -- the coefficient the pressure_matrix
    generator computes from the node spacings.  Interior-only rows sum to
    zero to rounding."""
    mg = pkg.meshgen
    case = mg.calcps_case((7, 6, 5), perturb=0.0, ratio=1.1,
                          bc={"xmin": "symmetry", "xmax": "symmetry", "ymin": "wall", "ymax": "wall",
                              "zmin": "symmetry", "zmax": "symmetry"})
    fields = case.fields()
    fields["asave"] = case.volume.copy()
    fields["asave"][case.volume == 0] = 1.0
    out = oracle.calcps(case, fields)
    ref = mg.pressure_matrix(mg.channel_mesh(7, 6, 5, ratio=1.1), pin_cell=0)
    got = case.matrix(out["aval"])
    A, B = got.to_scipy().tocsr(), ref.to_scipy().tocsr()
    off = A - B
    np.testing.assert_allclose(off.data, 0.0, atol=1e-13 * np.abs(B.data).max())
    # no boundary term touches the diagonal with these types: rows sum to 0
    rs = np.asarray(A.sum(axis=1)).ravel()
    assert np.all(np.abs(rs) <= 1e-13 * A.diagonal())


def test_invariants_perturbed_all_types(pkg, oracle):
    mg = pkg.meshgen
    bcs = {"xmin": "inflow", "xmax": "outflow", "ymin": "wall", "ymax": "convect", "zmin": "pressure",
           "zmax": "buffer"}
    case = mg.calcps_case((9, 8, 6), seed=11, bc=bcs, face_seed=5)
    out = oracle.calcps(case)
    aval = out["aval"]
    interior = case.sidec[2] > 0
    a1 = aval[case.sidaij[0, interior] - 1]
    a2 = aval[case.sidaij[1, interior] - 1]
    assert np.array_equal(a1, a2) and np.all(a1 < 0)
    # cells with no boundary face: zero row sums to rounding
    A = case.matrix(aval).to_scipy().tocsr()
    touched = np.zeros(case.nc, bool)
    touched[case.sidec[1, ~interior] - 1] = True
    rs = np.asarray(A.sum(axis=1)).ravel()
    assert np.all(np.abs(rs[~touched]) <= 1e-12 * A.diagonal()[~touched])
    # OUTFLOW / CONVECT / PRESSURE / BUFFER faces add to the diagonal
    assert np.all(rs[touched] >= -1e-12 * A.diagonal()[touched])
    # volume-weighted mean of b is removed (CalcPS.f90:201-204)
    V = case.volume[case.nbc + 1:]
    bv = out["b"] * V
    assert abs(bv.sum()) <= 1e-12 * np.abs(bv).sum()
    # BUFFER faces write Abou, the others leave it zero
    buf = case.typebc[:case.nbc] == mg.BC_CODES["buffer"]
    assert np.all(out["abou"][buf] < 0) and np.all(out["abou"][~buf] == 0)
    # SYMMETRY / WALL faces carry no flux
    t = case.typebc[case.sidec[2, ~interior] + case.nbc]
    wall = (t == mg.BC_CODES["wall"]) | (t == mg.BC_CODES["symmetry"])
    assert np.all(out["flux"][~interior][wall] == 0.0)


def test_face_order_independence(pkg, oracle):
    """Shuffling the faces changes only the order of the diagonal and b
    accumulations (rounding), nothing else."""
    mg = pkg.meshgen
    bcs = {"xmin": "inflow", "xmax": "outflow", "zmax": "pressure"}
    a = mg.calcps_case((8, 7, 5), seed=13, bc=bcs)
    b = mg.calcps_case((8, 7, 5), seed=13, bc=bcs, face_seed=21)
    oa, ob = oracle.calcps(a), oracle.calcps(b)
    Aa = a.matrix(oa["aval"]).to_scipy().tocsr()
    Ab = b.matrix(ob["aval"]).to_scipy().tocsr()
    np.testing.assert_allclose((Aa - Ab).data, 0.0, atol=1e-13 * np.abs(Aa.data).max())
    scale = np.abs(oa["b"]).max()
    np.testing.assert_allclose(oa["b"], ob["b"], rtol=0, atol=1e-12 * scale)
    assert oa["errmax"] == pytest.approx(ob["errmax"], rel=1e-13)


# ---------------------------------------------------------------------------
# the correction after the pressure solve (CorUVW.f90:36-121)
# ---------------------------------------------------------------------------
def test_coruvw_two_cell_hand_values(pkg, oracle):
    """PP = (1, 3) in the cells and 2 in the BUFFER cell -5.  Velocities:
    U1 = 2 - 1*1/2/1 = 1.5, U2 = 1 - 3*1/0.5/2 = -2; the PRESSURE cell -3
    (face 4, c1 = 2) takes U2.  Fluxes: face 1, 14 + (3-1)*(-2.5) = 9;
    face 6, -1 + (2-1)*(-1.25) = -2.25; the others keep their values.
    Mass error: cell 1, -9 - 2 - 0 + 2.25 = -8.75; cell 2, 9 - 2 - 2 = 5,
    minus its time term 2 = 3; divided by V = 1; errmax 8.75."""
    case = _two_cell_case(pkg)
    asm = oracle.calcps(case)
    nb = case.nbc
    pp = np.zeros(nb + case.nc + 1)
    pp[nb + 1], pp[nb + 2], pp[nb - 5] = 1.0, 3.0, 2.0
    out = oracle.coruvw(case, pp, asm["aval"], asm["abou"], asm["flux"])
    assert out["u"][nb + 1] == 1.5 and out["u"][nb + 2] == -2.0 and out["u"][nb - 3] == -2.0
    assert out["v"][nb + 1] == 0.5 and out["w"][nb - 3] == 0.0
    assert out["flux"].tolist() == [9.0, 2.0, 2.0, 2.0, 0.0, -2.25]
    assert out["b"].tolist() == [-8.75, 3.0]
    assert out["errmax"] == 8.75


def test_coruvw_closed_box_continuity(pkg, oracle):
    """Closed box (walls and symmetry): the pressure matrix has no boundary
    terms, so for any pressure correction PP the corrected fluxes leave
    b' V = b_pre - (A PP) in every cell, b_pre the assembly's source before
    its mean is removed (CalcPS.f90:188-199) -- the mass error that the
    correction does not remove.  A PP is taken through scipy, a path
    independent of both restatements."""
    mg = pkg.meshgen
    case = mg.calcps_case((8, 7, 6), seed=17, bc={"xmin": "wall", "xmax": "symmetry", "ymin": "wall",
                                                  "ymax": "wall", "zmin": "symmetry", "zmax": "symmetry"})
    asm = oracle.calcps(case)
    A = case.matrix(asm["aval"]).to_scipy()
    rng = np.random.default_rng(19)
    pp = np.zeros(case.nbc + case.nc + 1)
    pp[case.nbc + 1:] = rng.standard_normal(case.nc) * 1e-3
    out = oracle.coruvw(case, pp, asm["aval"], asm["abou"], asm["flux"])
    V = case.volume[case.nbc + 1:]
    want = asm["b_pre"] - A @ pp[case.nbc + 1:]
    scale = np.abs(asm["b_pre"]).max() + abs(A).max() * np.abs(pp).max() * 8
    np.testing.assert_allclose(out["b"] * V, want, rtol=0, atol=1e-13 * scale)


def _monitor_case(pkg):
    """The two-cell mesh set up for the CFL and Peclet monitor: material 1 on
    cell 1 and its BUFFER cell -5, material 2 on cell 2; f 0.25 at the
    interior face (RhoS = 0.25 * 1 + 0.75 * 3 = 2.5; 1.5 with the sides
    swapped) and 0.5 at the BUFFER face (RhoS = 0.5 * 1 + 0.5 * 3 = 2);
    Scoef 4 and 2 there, |D|^2 1 and 4; the four other boundary faces carry
    a flux of 1000 that the monitor must not see."""
    case = _two_cell_case(pkg)
    at = lambda c: c + case.nbc                                         # noqa: E731
    rho = case.rho_n.copy()
    rho[at(1)], rho[at(2)], rho[at(-5)] = 1.0, 3.0, 3.0
    f = case.f.copy()
    f[0], f[5] = 0.25, 0.5
    dx = case.dx.copy()
    dx[0], dx[5] = 1.0, -2.0
    material = np.ones(case.nbc + case.nc + 1, np.int32)
    material[at(2)], material[at(-2)], material[at(-3)] = 2, 2, 2
    scoef = np.array([4.0, 1.0, 1.0, 1.0, 1.0, 2.0])
    flux = np.array([8.0, 1000.0, -1000.0, 1000.0, 1000.0, -48.0])
    return dataclasses.replace(case, rho_n=rho, f=f, dx=dx), material, scoef, flux


def test_cfl_peclet_hand_values(pkg, oracle):
    """CorUVW.f90:124-147 with dt 0.5, VISc 0.25: interior face CFL |0.5 * 8 /
    (2.5 * 4 * 1)| = 0.4, Pe |8 / 4 / 0.25| = 8; BUFFER face CFL |0.5 * -48 /
    (2 * 2 * 4)| = 1.5, Pe |-48 / 2 / 0.25| = 96.  Material 1 sees both faces
    (cell 1 is c1 of each), material 2 only the interior one (its c2).  With
    VISc = 0 the Peclet number divides by TINY = 1e-64 and stays finite."""
    case, material, scoef, flux = _monitor_case(pkg)
    cfl, pe = oracle.cfl_peclet(case, material, 2, scoef, case.rho_n, flux, 0.5, 0.25)
    assert cfl.tolist() == [1.5, 0.4]
    assert pe.tolist() == [96.0, 8.0]
    cfl0, pe0 = oracle.cfl_peclet(case, material, 2, scoef, case.rho_n, flux, 0.5, 0.0)
    assert cfl0.tolist() == [1.5, 0.4]
    assert pe0.tolist() == [24.0 / 1e-64, 2.0 / 1e-64]
    # a material with no face of its own reads 0
    cfl3, pe3 = oracle.cfl_peclet(case, material, 3, scoef, case.rho_n, flux, 0.5, 0.25)
    assert cfl3.tolist() == [1.5, 0.4, 0.0] and pe3.tolist() == [96.0, 8.0, 0.0]
