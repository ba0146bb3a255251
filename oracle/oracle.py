"""ctypes front end of the two CPU checkers of the Krylov path.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product library.

* :class:`Oracle` -- the C restatement ``liboracle.so`` (krylov_oracle.c),
  which follows Process/CG.f90, BiCG.f90, CGS.f90, Prec1.f90, Resid.f90
  operation by operation; it runs wherever gcc built it, on any host.
* :class:`Reference` -- the reference's own Fortran solvers compiled from
  their sources into ``_ref/libtflows_ref.so`` by ``oracle/Makefile``.

Both expose the reference calling convention: ``cg(N, NB, NONZ, A, Acol,
Arow, Adia, Ab, x, r1, prec, niter, tol)`` updates ``x`` and ``r1`` in place
and returns ``(niter, IniRes, FinRes)``.  ``IniRes`` is returned as ``None``
when the solver did not write it (the early exit on ||b|| < tol,
Process/CG.f90:69-72).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libtflows_ref.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_SENTINEL = -1.2345e300


def build(with_reference: bool = True) -> None:
    """Compile the checkers (``make -C oracle``)."""
    target = "all" if with_reference else "liboracle.so"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


class _Solvers:
    _prefix = ""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
        self.lib = C.CDLL(path)
        ip, dp = C.POINTER(C.c_int32), C.POINTER(C.c_double)
        for name in ("cg", "bicg", "cgs"):
            f = getattr(self.lib, self._prefix + name + "_")
            f.restype = None
            f.argtypes = [ip, ip, ip, _f64p, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p,
                          ip, ip, dp, dp, dp]
        f = getattr(self.lib, self._prefix + "prec1_")
        f.restype = None
        f.argtypes = [ip, ip, _f64p, _i32p, _i32p, _i32p, _f64p, ip]
        f = getattr(self.lib, self._prefix + "resid_")
        f.restype = None
        f.argtypes = [ip, ip, ip, _f64p, _i32p, _i32p, _f64p, _f64p, _f64p]

    def _setup(self, n: int, nb: int) -> None:
        pass

    def solve(self, method: str, N, NB, NONZ, A, Acol, Arow, Adia, Ab, x, r1, prec, niter, tol):
        """Call ``method`` in {'cg','bicg','cgs'}; x (length NB+N+1, Fortran
        x(-NB:N)) and r1 (length N) are updated in place."""
        self._setup(N, NB)
        n_, nb_, nz_ = C.c_int32(N), C.c_int32(NB), C.c_int32(NONZ)
        pr_, it_ = C.c_int32(prec), C.c_int32(niter)
        tol_ = C.c_double(tol)
        ini, fin = C.c_double(_SENTINEL), C.c_double(0.0)
        Ab = Ab if Ab is not None and len(Ab) else np.zeros(1)
        assert x.dtype == np.float64 and len(x) == NB + N + 1
        getattr(self.lib, self._prefix + method + "_")(
            C.byref(n_), C.byref(nb_), C.byref(nz_), A, Acol, Arow, Adia, Ab, x, r1,
            C.byref(pr_), C.byref(it_), C.byref(tol_), C.byref(ini), C.byref(fin))
        return it_.value, (None if ini.value == _SENTINEL else ini.value), fin.value

    def cg(self, *a):
        return self.solve("cg", *a)

    def bicg(self, *a):
        return self.solve("bicg", *a)

    def cgs(self, *a):
        return self.solve("cgs", *a)

    def prec1(self, N, NONZ, A, Acol, Arow, Adia, prec):
        D = np.zeros(N)
        getattr(self.lib, self._prefix + "prec1_")(
            C.byref(C.c_int32(N)), C.byref(C.c_int32(NONZ)), A, Acol, Arow, Adia, D,
            C.byref(C.c_int32(prec)))
        return D

    def resid(self, N, NB, NONZ, A, Acol, Arow, Ab, x, r1):
        Ab = Ab if Ab is not None and len(Ab) else np.zeros(1)
        getattr(self.lib, self._prefix + "resid_")(
            C.byref(C.c_int32(N)), C.byref(C.c_int32(NB)), C.byref(C.c_int32(NONZ)),
            A, Acol, Arow, Ab, x, r1)

    # convenience: run a LinearSystem-like object; returns a result dict
    def run(self, method: str, m, b, x0=None, prec=1, niter=2000, tol=0.0, nb=0):
        x = np.zeros(nb + m.n + 1)
        if x0 is not None:
            x[nb + 1:] = x0
        r1 = np.array(b, dtype=np.float64, copy=True)
        it, ini, fin = self.solve(method, m.n, nb, m.nonz, m.aval, m.acol, m.arow, m.adia,
                                  np.zeros(max(nb, 1)), x, r1, prec, niter, tol)
        return dict(x=x[nb + 1:].copy(), r=r1, niter=it, ini_res=ini, fin_res=fin)


class Oracle(_Solvers):
    """C restatement (oracle/krylov_oracle.c)."""
    _prefix = "oracle_"

    def __init__(self, path: str = LIB_ORACLE):
        super().__init__(path)
        self.lib.oracle_get_error.restype = C.c_double
        self.lib.oracle_spmv.restype = None
        self.lib.oracle_spmv.argtypes = [C.c_int, _f64p, _i32p, _i32p, _f64p, _f64p]
        self.lib.oracle_prec_apply.restype = None
        self.lib.oracle_prec_apply.argtypes = [C.c_int, _f64p, _i32p, _i32p, _i32p, _f64p,
                                               C.c_int, _f64p, _f64p]

    def reset_state(self):
        self.lib.oracle_reset_state()

    def get_error(self) -> float:
        return self.lib.oracle_get_error()

    def set_error(self, e: float):
        self.lib.oracle_set_error(C.c_double(e))

    def spmv(self, m, p):
        q = np.zeros(m.n)
        self.lib.oracle_spmv(m.n, m.aval, m.acol, m.arow, np.ascontiguousarray(p, np.float64), q)
        return q

    def prec_apply(self, m, D, prec, u):
        z = np.zeros(m.n)
        self.lib.oracle_prec_apply(m.n, m.aval, m.acol, m.arow, m.adia, D, prec,
                                   np.ascontiguousarray(u, np.float64), z)
        return z

    def calcps(self, case, fields=None, nmat=None, dt=None):
        """Pressure-correction assembly (oracle/calcps_oracle.c, after
        Process/CalcPS.f90:51-202) of a meshgen.CalcPSCase; ``fields``
        overrides the case's per-call arrays.  Returns dict aval (NONZ),
        b (NC), flux (NS), abou (NbC, Abou(-NbC:-1)), errmax, mean, and
        b_pre (b before the mean is removed)."""
        f = dict(case.fields(), **(fields or {}))
        fn = self.lib.oracle_calcps
        if not getattr(fn, "_typed", False):
            fn.restype = None
            fn.argtypes = ([C.c_int] * 5 + [_i32p] * 4 + [_f64p] * 8 + [_f64p] * 11 + [C.c_double] +
                           [_f64p] * 4 + [C.POINTER(C.c_double)] * 2 + [_f64p])
            fn._typed = True
        aval = np.zeros(case.nonz)
        b = np.zeros(case.nc)
        flux = np.zeros(case.ns)
        abou = np.zeros(max(case.nbc, 1))
        em, mn = C.c_double(0.0), C.c_double(0.0)
        b_pre = np.zeros(case.nc)
        c32 = lambda a: np.ascontiguousarray(a, np.int32)   # noqa: E731
        c64 = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
        fn(case.nc, case.nbc, case.ns, case.nonz, case.nmat if nmat is None else nmat,
           c32(case.sidec.T), c32(case.sidaij.T), c32(case.adia), c32(case.typebc),
           *(c64(getattr(case, k)) for k in ("sx", "sy", "sz", "dx", "dy", "dz", "f", "volume")),
           *(c64(f[k]) for k in ("asave", "rho_n", "rho_o", "rho_oo", "u", "v", "w", "p", "px", "py", "pz")),
           case.dt if dt is None else dt, aval, b, flux, abou, C.byref(em), C.byref(mn), b_pre)
        return dict(aval=aval, b=b, flux=flux, abou=abou[:case.nbc], errmax=em.value, mean=mn.value, b_pre=b_pre)

    def coruvw(self, case, pp, aval, abou, flux, fields=None, dt=None):
        """The correction after the pressure solve (oracle/calcps_oracle.c,
        after Process/CorUVW.f90:36-121).  pp: (-NbC:NC); aval (NONZ); abou
        (NbC); flux (NS).  Returns dict u, v, w ((-NbC:NC)), flux, b (NC),
        errmax; the inputs are not modified."""
        f = dict(case.fields(), **(fields or {}))
        fn = self.lib.oracle_coruvw
        if not getattr(fn, "_typed", False):
            fn.restype = C.c_double
            fn.argtypes = [C.c_int] * 3 + [_i32p] * 3 + [_f64p] * 11 + [C.c_double] + [_f64p] * 5
            fn._typed = True
        c64 = lambda a: np.ascontiguousarray(a, np.float64).copy()   # noqa: E731
        u, v, w, fl = c64(f["u"]), c64(f["v"]), c64(f["w"]), c64(flux)
        b = np.zeros(case.nc)
        ab = np.zeros(max(case.nbc, 1))
        ab[:case.nbc] = abou
        em = fn(case.nc, case.nbc, case.ns, np.ascontiguousarray(case.sidec.T, np.int32),
                np.ascontiguousarray(case.sidaij.T, np.int32), np.ascontiguousarray(case.typebc, np.int32),
                c64(case.volume), *(c64(f[k]) for k in ("asave", "rho_n", "rho_o", "rho_oo", "px", "py", "pz")),
                c64(pp), c64(aval), ab, case.dt if dt is None else dt, u, v, w, fl, b)
        return dict(u=u, v=v, w=w, flux=fl, b=b, errmax=em)


    def newuvw(self, case, uf, params):
        """Momentum assembly of one component (oracle/calcps_oracle.c, after
        Process/NewUVW.f90:137-499) on a meshgen.CalcPSCase with the inputs
        of meshgen.uvw_fields and the flags of meshgen.uvw_params.  Returns
        dict aval (NONZ), b (NC), abou (NbC), asave ((-NbC:NC), cells 1..NC
        written for var = 1, zero elsewhere), rhoui ((-NbC:NC)); raises
        ValueError for a BLEND this restatement does not cover (NO next to
        YES).  uf["flux"], the mass flux whose sign picks ConvScheme's
        upwind cell, defaults to uf["flux_u"]."""
        fn = self.lib.oracle_newuvw
        if not getattr(fn, "_typed", False):
            fn.restype = C.c_int
            fn.argtypes = [C.c_int] * 9 + [_i32p] * 7 + [_f64p] * 28 + [C.c_double, C.c_int] + [C.c_double] * 2 + \
                [_f64p] * 7
            fn._typed = True
        var = int(params["var"])
        a = _uvw_args(case, uf, var)
        nall = case.nbc + case.nc + 1
        c32 = lambda x: np.ascontiguousarray(x, np.int32)     # noqa: E731
        c64 = lambda x: np.ascontiguousarray(x, np.float64)   # noqa: E731
        aval = np.zeros(case.nonz)
        b = np.zeros(case.nc)
        abou = np.zeros(max(case.nbc, 1))
        asave = np.zeros(nall)
        rhoui = np.zeros(nall)
        cbuf, xbuf = np.zeros(nall), np.zeros(nall)
        rc = fn(case.nc, case.nbc, case.ns, case.nonz, var, params["convec"], params["diffus"], params["cross"],
                params["inert"], c32(case.sidec.T), c32(case.sidaij.T), c32(case.adia), c32(case.typebc),
                c32(uf["material"]), c32(params["blend"]), c32(params["blend_tem"]), c64(params["urfc"]),
                c64(case.f), c64(uf["ff"]), c64(uf["scoef"]), c64(a["flux_u"]), c64(uf.get("flux", a["flux_u"])),
                *(c64(x) for x in a["S"]),
                *(c64(x) for x in a["D"]), c64(case.volume), c64(a["rho_n"]), c64(a["rho_o"]), c64(a["rho_oo"]),
                c64(a["ui_n"]), c64(a["ui_o"]), c64(a["ui_oo"]), c64(a["source"]), c64(a["pi"]),
                *(c64(x) for x in a["g"]), float(params["visc"]), int(params.get("visc_rho", 0)),
                float(params["dt"]), float(params["urf"]),
                aval, b, abou, asave[case.nbc + 1:], rhoui, cbuf, xbuf)
        if rc != 0:
            raise ValueError("BLEND not covered by the restatement (a ConvScheme face scheme)")
        return dict(aval=aval, b=b, abou=abou[:case.nbc], asave=asave, rhoui=rhoui)


def _calczmix(self, case, zf, params):
    """Mixture-fraction assembly (oracle/calcps_oracle.c, after
    Process/CalcZmix.f90:100-387) of a meshgen.CalcPSCase with the inputs of
    meshgen.zmix_fields; params as meshgen.uvw_params with BLEND_TEM in
    "blend_tem", URFC_Tem in "urfc", VIScZmix in "visc", Zmix%URF in "urf".
    Returns dict aval, b, abou (NbC), rhoz ((-NbC:NC)); zf["flux"] as for
    newuvw."""
    fn = self.lib.oracle_calczmix
    if not getattr(fn, "_typed", False):
        fn.restype = C.c_int
        fn.argtypes = [C.c_int] * 8 + [_i32p] * 6 + [_f64p] * 23 + [C.c_double, C.c_int] + [C.c_double] * 2 + \
            [_f64p] * 6
        fn._typed = True
    nall = case.nbc + case.nc + 1
    c32 = lambda x: np.ascontiguousarray(x, np.int32)     # noqa: E731
    c64 = lambda x: np.ascontiguousarray(x, np.float64)   # noqa: E731
    aval = np.zeros(case.nonz)
    b = np.zeros(case.nc)
    abou = np.zeros(max(case.nbc, 1))
    rhoz = np.zeros(nall)
    cbuf, xbuf = np.zeros(nall), np.zeros(nall)
    rc = fn(case.nc, case.nbc, case.ns, case.nonz, params["convec"], params["diffus"], params["cross"],
            params["inert"], c32(case.sidec.T), c32(case.sidaij.T), c32(case.adia), c32(case.typebc),
            c32(zf["material"]), c32(params["blend_tem"]), c64(params["urfc"]), c64(case.f), c64(zf["ff"]),
            c64(zf["scoef"]), c64(zf["flux_u"]), c64(zf.get("flux", zf["flux_u"])), c64(case.sx), c64(case.sy),
            c64(case.sz), c64(case.dx),
            c64(case.dy), c64(case.dz), c64(case.volume), c64(case.rho_n), c64(case.rho_o), c64(case.rho_oo),
            c64(zf["z"]), c64(zf["z_o"]), c64(zf["z_oo"]), c64(zf["src_z"]), c64(zf["zx"]), c64(zf["zy"]),
            c64(zf["zz"]), float(params["visc"]), int(params.get("visc_rho", 0)), float(params["dt"]),
            float(params["urf"]), aval, b, abou, rhoz, cbuf, xbuf)
    if rc != 0:
        raise ValueError("BLEND_TEM not covered by the restatement (a ConvScheme face scheme)")
    return dict(aval=aval, b=b, abou=abou[:case.nbc], rhoz=rhoz)


def _zmix_update(self, case, rhoz, rho_n, z_n):
    """CalcZmix.f90:401-405 on copies: returns the new z ((-NbC:NC))."""
    fn = self.lib.oracle_zmix_update
    if not getattr(fn, "_typed", False):
        fn.restype = None
        fn.argtypes = [C.c_int, C.c_int, _f64p, _f64p, _f64p]
        fn._typed = True
    z = np.array(z_n, dtype=np.float64, copy=True)
    fn(case.nc, case.nbc, np.ascontiguousarray(rhoz, np.float64), np.ascontiguousarray(rho_n, np.float64), z)
    return z


def _calminmax(self, case, phi):
    """Process/CalMinMax.f90 on phi ((-NbC:NC)): returns (PHImax, PHImin)."""
    fn = self.lib.oracle_calminmax
    if not getattr(fn, "_typed", False):
        fn.restype = None
        fn.argtypes = [C.c_int] * 3 + [_i32p] * 2 + [_f64p] * 3
        fn._typed = True
    mx, mn = np.zeros(len(phi)), np.zeros(len(phi))
    fn(case.nc, case.nbc, case.ns, np.ascontiguousarray(case.sidec.T, np.int32),
       np.ascontiguousarray(case.typebc, np.int32), np.ascontiguousarray(phi, np.float64), mx, mn)
    return mx, mn


def _cfl_peclet(self, case, material, nmat, scoef, rho_n, flux, dt, visc):
    """The CFL and Peclet monitor of CorUVW.f90:124-147 (oracle/calcps_oracle.c)
    on a meshgen.CalcPSCase: material ((-NbC:NC), 1-based), Scoef (NS),
    rho_n ((-NbC:NC)), flux (NS).  Returns (cflmax, pemax), Nmat values
    each."""
    fn = self.lib.oracle_cfl_peclet
    if not getattr(fn, "_typed", False):
        fn.restype = None
        fn.argtypes = [C.c_int] * 4 + [_i32p] * 3 + [_f64p] * 7 + [C.c_double] * 2 + [_f64p] * 2
        fn._typed = True
    c32 = lambda x: np.ascontiguousarray(x, np.int32)     # noqa: E731
    c64 = lambda x: np.ascontiguousarray(x, np.float64)   # noqa: E731
    cfl, pe = np.zeros(max(nmat, 1)), np.zeros(max(nmat, 1))
    fn(case.nc, case.nbc, case.ns, nmat, c32(case.sidec.T), c32(case.typebc), c32(material), c64(case.f),
       c64(scoef), c64(case.dx), c64(case.dy), c64(case.dz), c64(rho_n), c64(flux), float(dt), float(visc), cfl, pe)
    return cfl[:nmat], pe[:nmat]


Oracle.cfl_peclet = _cfl_peclet
Oracle.calczmix = _calczmix
Oracle.calminmax = _calminmax
Oracle.zmix_update = _zmix_update


def _uvw_args(case, uf, var):
    """newuvw's arrays for var, as Process/Processor.f90:314-340 passes them."""
    grads = {1: ("ux", "uy", "uz", "vx", "wx", "vy", "wz"), 2: ("vy", "vx", "vz", "uy", "wy", "ux", "wz"),
             3: ("wz", "wx", "wy", "uz", "vz", "ux", "vy")}[var]
    order = {1: (0, 1, 2), 2: (1, 0, 2), 3: (2, 0, 1)}[var]
    comp = "uvw"[var - 1]
    S = (case.sx, case.sy, case.sz)
    D = (case.dx, case.dy, case.dz)
    return {"rho_n": case.rho_n, "rho_o": case.rho_o, "rho_oo": case.rho_oo,
            "ui_n": uf.get(comp, getattr(case, comp)), "ui_o": uf[comp + "_o"], "ui_oo": uf[comp + "_oo"],
            "source": uf["src_" + comp], "pi": getattr(case, "p" + "xyz"[var - 1]), "g": [uf[k] for k in grads],
            "S": [S[q] for q in order], "D": [D[q] for q in order], "flux_u": uf["flux_u"]}


class Reference(_Solvers):
    """The reference Fortran solvers, compiled from their own sources."""
    _prefix = ""

    def __init__(self, path: str = LIB_REF):
        super().__init__(path)
        self.lib.ref_setup_.restype = None
        self.lib.ref_setup_.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        self._shape = None

    def _setup(self, n, nb):
        if self._shape != (n, nb):
            self.lib.ref_setup_(C.byref(C.c_int32(n)), C.byref(C.c_int32(nb)))
            self._shape = (n, nb)

    def set_error(self, e: float):
        self.lib.ref_set_error_(C.byref(C.c_double(e)))

    def get_error(self) -> float:
        e = C.c_double(0.0)
        self.lib.ref_get_error_(C.byref(e))
        return e.value


def available_reference() -> bool:
    return os.path.exists(LIB_REF)
