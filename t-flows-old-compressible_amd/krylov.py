"""Host-side mirror of the reference solver interface over libtflows_krylov.so.

The reference subroutines (Process/CG.f90:2-3, BiCG.f90:2-3, CGS.f90:2-3,
Prec1.f90:2, Resid.f90:2) are called here with the same names, argument
order and meaning:

    niter, IniRes, FinRes = cg(N, NB, NONZ, A, Acol, Arow, Adia, Ab, x, r1,
                               prec, niter, tol)

``x`` (length NB+N+1, Fortran x(-NB:N)) and ``r1`` (length N) are numpy
arrays updated in place exactly as the Fortran dummies are; the three
scalar outputs are returned.  ``IniRes`` is ``None`` when the solver left it
unwritten (the ||b|| < tol early exit, Process/CG.f90:69-72).

The computation always runs on the GPU through the C ABI; if the shared
library is missing or no GPU is present the calls raise -- there is no CPU
fallback.

:class:`DeviceSystem` wraps the device-resident API for callers that keep
the matrix and vectors in HBM (torch tensors on ``cuda``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtflows_krylov.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "tflows_krylov.h")

METHODS = {"cg": 0, "bicg": 1, "cgs": 2}

_ip = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_SENTINEL = -1.2345e300

_lib = None


class TFlowsError(RuntimeError):
    pass


def library_path() -> str:
    return LIB_PATH


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libtflows_krylov.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TFlowsError(f"{path} is not built; run __graft_entry__.build()")
    # torch bundles its own libamdhip64.so.7 with the same SONAME as the
    # system one; whichever loads first serves the whole process.  Load
    # torch's first so that torch tensors and this library share one HIP
    # runtime (the other order leaves torch without a usable device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name in ("cg_", "bicg_", "cgs_"):
        f = getattr(lib, name)
        f.restype = None
        f.argtypes = [_ip, _ip, _ip, _f64, _i32, _i32, _i32, _f64, _f64, _f64, _ip, _ip, _dp, _dp, _dp]
    lib.prec1_.restype = None
    lib.prec1_.argtypes = [_ip, _ip, _f64, _i32, _i32, _i32, _f64, _ip]
    lib.resid_.restype = None
    lib.resid_.argtypes = [_ip, _ip, _ip, _f64, _i32, _i32, _f64, _f64, _f64]
    lib.tflows_last_error.restype = C.c_char_p
    lib.tflows_version.restype = C.c_char_p
    vp = C.c_void_p
    lib.tflows_system_create.restype = C.c_int
    lib.tflows_system_create.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, C.c_int32, vp, C.POINTER(vp)]
    lib.tflows_system_destroy.argtypes = [vp]
    lib.tflows_system_set_values.argtypes = [vp, vp, C.c_int32, vp]
    lib.tflows_solve.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, _ip, C.c_double, _dp, _dp, vp]
    lib.tflows_spmv.argtypes = [vp, vp, vp, vp]
    lib.tflows_spmv32.argtypes = [vp, vp, vp, vp]
    lib.tflows_prec1.argtypes = [vp, C.c_int32, vp, vp]
    lib.tflows_resid.argtypes = [vp, vp, vp, vp]
    lib.tflows_system_levels.restype = C.c_int32
    lib.tflows_system_levels.argtypes = [vp]
    lib.tflows_dropin_stats.restype = C.c_int
    lib.tflows_dropin_stats.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.tflows_dropin_release.restype = C.c_int
    lib.tflows_dropin_release.argtypes = []
    lib.tflows_system_dic_sweep.restype = C.c_int32
    lib.tflows_system_dic_sweep.argtypes = [vp]
    _lib = lib
    _bind_extensions(lib)
    return lib


def _bind_extensions(lib):
    """Optional entry points of the mixed-precision, batched and adjoint
    components (bound when present in the library)."""
    vp = C.c_void_p
    if hasattr(lib, "tflows_solve_mixed"):
        f = lib.tflows_solve_mixed
        f.restype = C.c_int
        f.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, _ip, C.c_double, C.c_double, C.c_int32,
                      _dp, _dp, _ip, vp]
    for name in ("tflows_solve_adjoint",):
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = C.c_int
            f.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, _ip, C.c_double, _dp, _dp, vp]
    if hasattr(lib, "tflows_outer_grad"):
        lib.tflows_outer_grad.restype = C.c_int
        lib.tflows_outer_grad.argtypes = [vp, vp, vp, vp, vp]
        lib.tflows_transpose_values.restype = C.c_int
        lib.tflows_transpose_values.argtypes = [vp, vp, vp]
    if hasattr(lib, "tflows_solve_mixed_reliable"):
        f = lib.tflows_solve_mixed_reliable
        f.restype = C.c_int
        f.argtypes = [vp, C.c_int32, vp, vp, vp, _ip, C.c_double, C.c_int32, _dp, _dp, _ip, vp]
    if hasattr(lib, "tflows_system_set_lanes"):
        lib.tflows_system_set_lanes.restype = C.c_int
        lib.tflows_system_set_lanes.argtypes = [vp, C.c_int32]
    if hasattr(lib, "tflows_system_reordered"):
        lib.tflows_system_reordered.restype = C.c_int32
        lib.tflows_system_reordered.argtypes = [vp]
    for name in ("tflows_system_fp32_layout", "tflows_system_fp64_layout"):
        if hasattr(lib, name):
            getattr(lib, name).restype = C.c_int
            getattr(lib, name).argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), vp]
    if hasattr(lib, "tflows_batch_create"):
        pp = C.POINTER(vp)
        lib.tflows_batch_create.restype = C.c_int
        lib.tflows_batch_create.argtypes = [C.c_int32, vp, vp, pp, pp, pp, pp]
        lib.tflows_batch_destroy.restype = C.c_int
        lib.tflows_batch_destroy.argtypes = [vp]
        lib.tflows_batch_set_values.restype = C.c_int
        lib.tflows_batch_set_values.argtypes = [vp, C.c_int32, vp, C.c_int32, vp]
        lib.tflows_batch_layout.restype = C.c_int
        lib.tflows_batch_layout.argtypes = [vp, _ip, C.POINTER(C.c_int64), vp]
        lib.tflows_batch_configure.restype = C.c_int
        lib.tflows_batch_configure.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        lib.tflows_batch_config.restype = C.c_int
        lib.tflows_batch_config.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, _ip, _ip, _ip, _ip]
        lib.tflows_batch_set_form.restype = C.c_int
        lib.tflows_batch_set_form.argtypes = [vp, C.c_int32]
        lib.tflows_batch_form.restype = C.c_int32
        lib.tflows_batch_form.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
        lib.tflows_batch_set_groups.restype = C.c_int
        lib.tflows_batch_set_groups.argtypes = [vp, C.c_int32]
        lib.tflows_batch_groups.restype = C.c_int32
        lib.tflows_batch_groups.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
        lib.tflows_batch_solve.restype = C.c_int
        lib.tflows_batch_solve.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp]
        lib.tflows_batch_solve_mixed.restype = C.c_int
        lib.tflows_batch_solve_mixed.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, C.c_double,
                                                 C.c_int32, vp, vp, vp, _ip, vp]
    if hasattr(lib, "tflows_mesh_load"):
        lib.tflows_mesh_load.restype = C.c_int
        lib.tflows_mesh_load.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(vp)]
        lib.tflows_mesh_destroy.restype = C.c_int
        lib.tflows_mesh_destroy.argtypes = [vp]
        lib.tflows_mesh_info.restype = C.c_int
        lib.tflows_mesh_info.argtypes = [vp] + [C.POINTER(C.c_int32)] * 6
        for name in ("tflows_mesh_int_array", "tflows_mesh_real_array"):
            getattr(lib, name).restype = C.c_int
            getattr(lib, name).argtypes = [vp, C.c_char_p, vp, C.c_int64]
        lib.tflows_mesh_typebc.restype = C.c_int
        lib.tflows_mesh_typebc.argtypes = [vp, vp, C.c_int32, vp, C.c_int64]
        lib.tflows_mesh_ff.restype = C.c_int
        lib.tflows_mesh_ff.argtypes = [vp, vp, vp, vp, C.c_int64]
    if hasattr(lib, "tflows_bfile_load"):
        lib.tflows_bfile_load.restype = C.c_int
        lib.tflows_bfile_load.argtypes = [C.c_char_p, C.c_int32, C.POINTER(vp)]
        lib.tflows_bfile_destroy.restype = C.c_int
        lib.tflows_bfile_destroy.argtypes = [vp]
        lib.tflows_bfile_info.restype = C.c_int
        lib.tflows_bfile_info.argtypes = [vp] + [C.POINTER(C.c_int32)] * 3 + [C.POINTER(C.c_double)]
        lib.tflows_bfile_materials.restype = C.c_int
        lib.tflows_bfile_materials.argtypes = [vp] * 5
        lib.tflows_bfile_bounds.restype = C.c_int
        lib.tflows_bfile_bounds.argtypes = [vp] * 8
        lib.tflows_bfile_inits.restype = C.c_int
        lib.tflows_bfile_inits.argtypes = [vp] * 6
        lib.tflows_bfile_name.restype = C.c_int
        lib.tflows_bfile_name.argtypes = [vp, C.c_int32, C.c_int32, C.c_char_p, C.c_int32]
    if hasattr(lib, "tflows_calcps_create"):
        lib.tflows_calcps_create.restype = C.c_int
        lib.tflows_calcps_create.argtypes = [C.c_int32] * 4 + [vp] * 12 + [C.c_int32, vp, C.POINTER(vp)]
        lib.tflows_calcps_destroy.restype = C.c_int
        lib.tflows_calcps_destroy.argtypes = [vp]
        lib.tflows_calcps_visit_order.restype = C.c_int32
        lib.tflows_calcps_visit_order.argtypes = [vp]
        lib.tflows_calcps_assemble.restype = C.c_int
        lib.tflows_calcps_assemble.argtypes = ([vp, C.c_int32, C.c_double] + [vp] * 11 + [vp] * 4 +
                                               [C.POINTER(C.c_double)] * 2 + [C.c_int32, vp])
        lib.tflows_calcps_uvw_setup.restype = C.c_int
        lib.tflows_calcps_uvw_setup.argtypes = [vp, vp, vp, vp, C.c_int32, C.c_int32, vp]
        lib.tflows_calcps_uvw_assemble.restype = C.c_int
        lib.tflows_calcps_uvw_assemble.argtypes = [vp, C.POINTER(UvwParams)] + [vp] * 17 + [vp] * 5 + [C.c_int32, vp]
        lib.tflows_calcps_zmix_assemble.restype = C.c_int
        lib.tflows_calcps_zmix_assemble.argtypes = [vp, C.POINTER(UvwParams)] + [vp] * 12 + [vp] * 4 + [C.c_int32, vp]
        lib.tflows_calcps_zmix_update.restype = C.c_int
        lib.tflows_calcps_zmix_update.argtypes = [vp, vp, vp, vp, C.c_int32, vp]
        lib.tflows_calcps_uvw_update.restype = C.c_int
        lib.tflows_calcps_uvw_update.argtypes = [vp, vp, vp, vp, C.c_int32, vp]
        lib.tflows_calcps_cfl_peclet.restype = C.c_int
        lib.tflows_calcps_cfl_peclet.argtypes = [vp, C.c_double, C.c_double, vp, vp, vp, vp, C.c_int32, vp]
        lib.tflows_calcps_correct.restype = C.c_int
        lib.tflows_calcps_correct.argtypes = ([vp, C.c_double] + [vp] * 10 + [vp] * 5 + [C.POINTER(C.c_double)] +
                                              [C.c_int32, vp])


def _check(rc: int):
    if rc != 0:
        raise TFlowsError(load_library().tflows_last_error().decode())


def declared_symbols(header: str = HEADER) -> list[str]:
    """Function names declared in include/tflows_krylov.h."""
    import re
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt)) -
                  {"if", "defined", "extern", "sizeof"})


# --------------------------------------------------------------------------
# drop-in interface (host arrays, reference argument lists)
# --------------------------------------------------------------------------
def _solve(name, N, NB, NONZ, A, Acol, Arow, Adia, Ab, x, r1, prec, niter, tol):
    lib = load_library()
    for arr, dt in ((A, np.float64), (x, np.float64), (r1, np.float64), (Acol, np.int32), (Arow, np.int32),
                    (Adia, np.int32)):
        if arr.dtype != dt or not arr.flags.c_contiguous:
            raise TypeError("arrays must be C-contiguous float64 / int32")
    if len(x) != NB + N + 1 or len(r1) != N or len(Acol) != N + 1 or len(A) < NONZ:
        raise ValueError("array lengths do not match N, NB, NONZ")
    Ab = Ab if Ab is not None and len(Ab) else np.zeros(1)
    it = C.c_int32(niter)
    ini, fin = C.c_double(_SENTINEL), C.c_double(0.0)
    getattr(lib, name)(C.byref(C.c_int32(N)), C.byref(C.c_int32(NB)), C.byref(C.c_int32(NONZ)), A, Acol, Arow,
                       Adia, np.ascontiguousarray(Ab, np.float64), x, r1, C.byref(C.c_int32(prec)), C.byref(it),
                       C.byref(C.c_double(tol)), C.byref(ini), C.byref(fin))
    return it.value, (None if ini.value == _SENTINEL else ini.value), fin.value


def cg(N, NB, NONZ, A, Acol, Arow, Adia, Ab, x, r1, prec, niter, tol):
    """GPU replacement of SUBROUTINE cg (Process/CG.f90:2)."""
    return _solve("cg_", N, NB, NONZ, A, Acol, Arow, Adia, Ab, x, r1, prec, niter, tol)


def bicg(N, NB, NONZ, A, Acol, Arow, Adia, Ab, x, r1, prec, niter, tol):
    """GPU replacement of SUBROUTINE bicg (Process/BiCG.f90:2)."""
    return _solve("bicg_", N, NB, NONZ, A, Acol, Arow, Adia, Ab, x, r1, prec, niter, tol)


def cgs(N, NB, NONZ, A, Acol, Arow, Adia, Ab, x, r1, prec, niter, tol):
    """GPU replacement of SUBROUTINE CGS (Process/CGS.f90:2)."""
    return _solve("cgs_", N, NB, NONZ, A, Acol, Arow, Adia, Ab, x, r1, prec, niter, tol)


def prec1(N, NONZ, A, Acol, Arow, Adia, prec):
    """GPU replacement of SUBROUTINE Prec1 (Process/Prec1.f90:2); returns D."""
    lib = load_library()
    D = np.zeros(N)
    lib.prec1_(C.byref(C.c_int32(N)), C.byref(C.c_int32(NONZ)), A, Acol, Arow, Adia, D, C.byref(C.c_int32(prec)))
    return D


def resid(N, NB, NONZ, A, Acol, Arow, Ab, x, r1):
    """GPU replacement of SUBROUTINE Resid (Process/Resid.f90:2); r1 in place."""
    lib = load_library()
    Ab = Ab if Ab is not None and len(Ab) else np.zeros(1)
    lib.resid_(C.byref(C.c_int32(N)), C.byref(C.c_int32(NB)), C.byref(C.c_int32(NONZ)), A, Acol, Arow,
               np.ascontiguousarray(Ab, np.float64), x, r1)


def dropin_stats() -> dict:
    """Device pattern builds and calls of the drop-ins so far: a call on the
    pattern of the one before uploads only values and vectors."""
    b, c = C.c_int64(0), C.c_int64(0)
    _check(load_library().tflows_dropin_stats(C.byref(b), C.byref(c)))
    return {"pattern_builds": int(b.value), "calls": int(c.value)}


def dropin_release():
    """Free the drop-ins' cached device pattern."""
    _check(load_library().tflows_dropin_release())


def solve_system(method: str, m, b, x0=None, prec=1, niter=2000, tol=0.0, nb=0):
    """Convenience wrapper over the drop-in for a meshgen.TFlowSMatrix."""
    x = np.zeros(nb + m.n + 1)
    if x0 is not None:
        x[nb + 1:] = x0
    r1 = np.array(b, dtype=np.float64, copy=True)
    it, ini, fin = _solve(method + "_", m.n, nb, m.nonz, m.aval, m.acol, m.arow, m.adia, np.zeros(max(nb, 1)),
                          x, r1, prec, niter, tol)
    return dict(x=x[nb + 1:].copy(), r=r1, niter=it, ini_res=ini, fin_res=fin)


# --------------------------------------------------------------------------
# device-resident interface (torch tensors in HBM)
# --------------------------------------------------------------------------
def _layout(code: int, entries: int) -> dict:
    """Operator layout from the library's code: sliced or CSR; bytes per
    stored column index (4, 2 for 16-bit offsets, 0 when the slices share
    their column offsets); whether only the non-negative offsets' slots are
    stored, lower entries read from their mirrors; entries held, padding
    included."""
    return {"sliced": code > 0, "shared": code in (3, 4), "symmetric": code == 4,
            "col_bytes": {3: 0, 4: 0, 2: 2}.get(code, 4), "stored_entries": int(entries)}


def _ptr(t) -> int:
    return int(t.data_ptr())


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(int(s.cuda_stream))


class DeviceSystem:
    """A matrix resident on the GPU (pattern + FP64 values), solved in place
    on device vectors.  ``matrix`` is a meshgen.TFlowSMatrix (host arrays)."""

    def __init__(self, matrix, stream=None):
        import torch
        self.lib = load_library()
        self.n, self.nonz = matrix.n, matrix.nonz
        self.matrix = matrix
        h = C.c_void_p()
        dev = torch.device("cuda")
        self._acol = torch.from_numpy(matrix.acol).to(dev)
        self._arow = torch.from_numpy(matrix.arow).to(dev)
        self._adia = torch.from_numpy(matrix.adia).to(dev)
        self.values = torch.from_numpy(matrix.aval).to(dev)
        torch.cuda.synchronize()
        _check(self.lib.tflows_system_create(self.n, self.nonz, _ptr(self._acol), _ptr(self._arow),
                                             _ptr(self._adia), 1, _stream(stream), C.byref(h)))
        self.handle = h
        self.set_values(self.values, stream)

    @classmethod
    def from_device(cls, n, nonz, acol, arow, adia, aval, stream=None):
        """A system from cuda tensors already in the T-FlowS layout (int32
        acol (N+1), arow and float64 aval (NONZ), int32 adia (N))."""
        self = cls.__new__(cls)
        self.lib = load_library()
        self.n, self.nonz, self.matrix = int(n), int(nonz), None
        self._acol, self._arow, self._adia, self.values = acol, arow, adia, aval
        h = C.c_void_p()
        _check(self.lib.tflows_system_create(self.n, self.nonz, _ptr(acol), _ptr(arow), _ptr(adia), 1,
                                             _stream(stream), C.byref(h)))
        self.handle = h
        self.set_values(aval, stream)
        return self

    def set_values(self, aval, stream=None):
        """aval: cuda float64 tensor of length NONZ (or nnz)."""
        _check(self.lib.tflows_system_set_values(self.handle, C.c_void_p(_ptr(aval)), 1, _stream(stream)))

    def solve(self, method: str, prec: int, x, r1, niter: int, tol: float, stream=None):
        """x, r1: cuda float64 tensors of length N, updated in place."""
        it = C.c_int32(niter)
        ini, fin = C.c_double(_SENTINEL), C.c_double(0.0)
        _check(self.lib.tflows_solve(self.handle, METHODS[method], prec, C.c_void_p(_ptr(x)),
                                     C.c_void_p(_ptr(r1)), C.byref(it), tol, C.byref(ini), C.byref(fin),
                                     _stream(stream)))
        return it.value, (None if ini.value == _SENTINEL else ini.value), fin.value

    def solve_mixed(self, method: str, prec: int, x, b, niter: int, tol: float, inner_rtol: float = 1e-4,
                    max_outer: int = 100, r_out=None, stream=None):
        """FP32 Krylov inside FP64 iterative refinement.  x: cuda float64
        tensor (guess in, solution out); b: cuda float64 tensor (unchanged).
        Returns (niter_total_inner, IniRes, FinRes, outer_steps) with the
        residual norms those of the FP64 true residual b - A x."""
        it = C.c_int32(niter)
        ini, fin = C.c_double(0.0), C.c_double(0.0)
        outer = C.c_int32(0)
        _check(self.lib.tflows_solve_mixed(self.handle, METHODS[method], prec, C.c_void_p(_ptr(x)),
                                           C.c_void_p(_ptr(b)),
                                           C.c_void_p(_ptr(r_out) if r_out is not None else 0), C.byref(it),
                                           tol, inner_rtol, max_outer, C.byref(ini), C.byref(fin),
                                           C.byref(outer), _stream(stream)))
        return it.value, ini.value, fin.value, outer.value

    def solve_mixed_reliable(self, prec: int, x, b, niter: int, tol: float, interval: int = 50, r_out=None,
                             stream=None):
        """CG in FP32 with FP64 residual replacement every `interval`
        iterations; returns (niter, IniRes, FinRes, replacements)."""
        it = C.c_int32(niter)
        ini, fin = C.c_double(0.0), C.c_double(0.0)
        rep = C.c_int32(0)
        _check(self.lib.tflows_solve_mixed_reliable(
            self.handle, prec, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(b)),
            C.c_void_p(_ptr(r_out) if r_out is not None else 0), C.byref(it), tol, interval, C.byref(ini),
            C.byref(fin), C.byref(rep), _stream(stream)))
        return it.value, ini.value, fin.value, rep.value

    def solve_adjoint(self, method: str, prec: int, lam, r1, niter: int, tol: float, stream=None):
        """A^T lam = r1 (r1 in: right-hand side, out: final residual) with the
        method and preconditioner of the forward solve."""
        it = C.c_int32(niter)
        ini, fin = C.c_double(_SENTINEL), C.c_double(0.0)
        _check(self.lib.tflows_solve_adjoint(self.handle, METHODS[method], prec, C.c_void_p(_ptr(lam)),
                                             C.c_void_p(_ptr(r1)), C.byref(it), tol, C.byref(ini), C.byref(fin),
                                             _stream(stream)))
        return it.value, (None if ini.value == _SENTINEL else ini.value), fin.value

    def outer_grad(self, lam, x, grad, stream=None):
        """grad[p] = -lam[i] * x[j] for every stored entry p = (i, j)."""
        _check(self.lib.tflows_outer_grad(self.handle, C.c_void_p(_ptr(lam)), C.c_void_p(_ptr(x)),
                                          C.c_void_p(_ptr(grad)), _stream(stream)))

    def transpose_values(self, out, stream=None):
        _check(self.lib.tflows_transpose_values(self.handle, C.c_void_p(_ptr(out)), _stream(stream)))

    def spmv(self, x, y, stream=None):
        _check(self.lib.tflows_spmv(self.handle, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(y)), _stream(stream)))

    def spmv32(self, x, y, stream=None):
        _check(self.lib.tflows_spmv32(self.handle, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(y)), _stream(stream)))

    def prec1(self, prec: int, d, stream=None):
        _check(self.lib.tflows_prec1(self.handle, prec, C.c_void_p(_ptr(d)), _stream(stream)))

    def resid(self, x, r, stream=None):
        _check(self.lib.tflows_resid(self.handle, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(r)), _stream(stream)))

    def levels(self) -> int:
        return int(self.lib.tflows_system_levels(self.handle))

    def dic_sweep(self) -> int:
        """Form of the DIC sweeps (3 one launch per triangle over compact
        tiles, 2 one launch per triangle row by row, 1 one launch per level
        over the level-major copy, 0 over the matrix rows)."""
        v = int(self.lib.tflows_system_dic_sweep(self.handle))
        if v < 0:
            _check(v)
        return v

    @property
    def reordered(self) -> bool:
        """True when the matrix is stored in a locality order internally."""
        return bool(self.lib.tflows_system_reordered(self.handle))

    def fp32_layout(self, stream=None) -> dict:
        """Layout of the FP32 inner operator: {"sliced": bool, "shared":
        bool (column offsets shared by each slice), "col_bytes": 4, 2 for
        16-bit column offsets, 0 when shared, "stored_entries": off-diagonal
        entries held, padding included}."""
        sl, ent = C.c_int32(0), C.c_int64(0)
        _check(self.lib.tflows_system_fp32_layout(self.handle, C.byref(sl), C.byref(ent), _stream(stream)))
        return _layout(sl.value, ent.value)

    def fp64_layout(self, stream=None) -> dict:
        """Layout of the FP64 operator of the iterations, as fp32_layout."""
        sl, ent = C.c_int32(0), C.c_int64(0)
        _check(self.lib.tflows_system_fp64_layout(self.handle, C.byref(sl), C.byref(ent), _stream(stream)))
        return _layout(sl.value, ent.value)

    def set_lanes(self, lanes: int):
        """SpMV form: lanes per row (1, 2, 4 or 8); results are identical."""
        _check(self.lib.tflows_system_set_lanes(self.handle, lanes))

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.tflows_system_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBatch:
    """K independent systems solved together on the GPU.

    ``matrices`` is a list of meshgen.TFlowSMatrix (any sizes and
    patterns).  Vectors are cuda float64 tensors of ``rows`` entries holding
    the systems back to back; ``offsets[k]`` is the first row of system k.
    """

    def __init__(self, matrices, stream=None):
        self.lib = load_library()
        self.k = len(matrices)
        self.matrices = list(matrices)
        n = np.array([m.n for m in matrices], dtype=np.int32)
        nonz = np.array([m.nonz for m in matrices], dtype=np.int32)
        keep = [(np.ascontiguousarray(m.acol, np.int32), np.ascontiguousarray(m.arow, np.int32),
                 np.ascontiguousarray(m.adia, np.int32)) for m in matrices]
        arr = C.c_void_p * self.k
        acol = arr(*[a[0].ctypes.data for a in keep])
        arow = arr(*[a[1].ctypes.data for a in keep])
        adia = arr(*[a[2].ctypes.data for a in keep])
        h = C.c_void_p()
        _check(self.lib.tflows_batch_create(self.k, n.ctypes.data, nonz.ctypes.data,
                                            C.cast(acol, C.POINTER(C.c_void_p)),
                                            C.cast(arow, C.POINTER(C.c_void_p)),
                                            C.cast(adia, C.POINTER(C.c_void_p)), C.byref(h)))
        self.handle = h
        rows = C.c_int32(0)
        ents = C.c_int64(0)
        offs = np.zeros(self.k + 1, dtype=np.int32)
        _check(self.lib.tflows_batch_layout(h, C.byref(rows), C.byref(ents), offs.ctypes.data))
        self.rows, self.entries, self.offsets = rows.value, ents.value, offs
        for k, m in enumerate(matrices):
            self.set_values(k, np.ascontiguousarray(m.aval[:m.nnz], np.float64))

    def set_values(self, k: int, aval, stream=None):
        """Values of system k (k = -1: all systems back to back); a numpy
        array (host) or a cuda float64 tensor (device)."""
        if isinstance(aval, np.ndarray):
            a = np.ascontiguousarray(aval, np.float64)
            _check(self.lib.tflows_batch_set_values(self.handle, k, a.ctypes.data, 0, None))
        else:
            _check(self.lib.tflows_batch_set_values(self.handle, k, C.c_void_p(_ptr(aval)), 1, _stream(stream)))

    def configure(self, grid: int = 0, lanes: int = 0, chunk: int = 0, rebalance: bool = True):
        _check(self.lib.tflows_batch_configure(self.handle, grid, lanes, chunk, 1 if rebalance else 0))

    def set_form(self, form: int):
        """0: tuned; 1: one launch per kernel; 2: one persistent launch per
        chunk of iterations (CG / BiCG with prec 0 or 1)."""
        _check(self.lib.tflows_batch_set_form(self.handle, form))

    def set_groups(self, mode: int):
        """0: tuned; 1: every system from the first iteration; 2: systems in
        groups that fit the MALL, one group after another."""
        _check(self.lib.tflows_batch_set_groups(self.handle, mode))

    def config(self, method: str = "cg", prec: int = 1, bits: int = 64) -> dict:
        g, l, c, r = (C.c_int32(0) for _ in range(4))
        _check(self.lib.tflows_batch_config(self.handle, bits, METHODS[method], prec, C.byref(g), C.byref(l),
                                            C.byref(c), C.byref(r)))
        form = int(self.lib.tflows_batch_form(self.handle, bits, METHODS[method], prec))
        groups = int(self.lib.tflows_batch_groups(self.handle, bits, METHODS[method], prec))
        return dict(grid=g.value, lanes=l.value, chunk=c.value, rebalances=r.value, form=form, groups=groups)

    def _arrays(self, cap, tol, fin_in):
        cap = np.ascontiguousarray(np.broadcast_to(np.asarray(cap, np.int32), (self.k,)))
        tol = np.ascontiguousarray(np.broadcast_to(np.asarray(tol, np.float64), (self.k,)))
        fin = np.ascontiguousarray(np.broadcast_to(np.asarray(fin_in, np.float64), (self.k,))).copy()
        ini = np.full(self.k, _SENTINEL)
        nit = np.zeros(self.k, np.int32)
        return cap, tol, fin, ini, nit

    def solve(self, method: str, prec: int, x, r1, cap, tol, fin_in=0.0, stream=None):
        """FP64 solve of every system; x, r1 cuda float64 tensors (r1 in: b,
        out: final residual).  cap, tol: scalars or per-system arrays.
        Returns dict of per-system arrays niter, ini_res (nan where not
        written), fin_res."""
        cap, tol, fin, ini, nit = self._arrays(cap, tol, fin_in)
        _check(self.lib.tflows_batch_solve(self.handle, METHODS[method], prec, C.c_void_p(_ptr(x)),
                                           C.c_void_p(_ptr(r1)), cap.ctypes.data, tol.ctypes.data,
                                           nit.ctypes.data, ini.ctypes.data, fin.ctypes.data, _stream(stream)))
        ini = np.where(ini == _SENTINEL, np.nan, ini)
        return dict(niter=nit, ini_res=ini, fin_res=fin)

    def solve_mixed(self, method: str, prec: int, x, b, cap, tol, inner_rtol: float = 1e-4,
                    max_outer: int = 100, r_out=None, stream=None):
        """FP32 inner solves inside FP64 refinement, per system; b unchanged.
        Returns dict niter (total inner iterations), ini_res, fin_res (FP64
        true residual norms), outer (refinement steps of the batch)."""
        cap, tol, fin, ini, nit = self._arrays(cap, tol, 0.0)
        outer = C.c_int32(0)
        _check(self.lib.tflows_batch_solve_mixed(
            self.handle, METHODS[method], prec, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(b)),
            C.c_void_p(_ptr(r_out) if r_out is not None else 0), cap.ctypes.data, tol.ctypes.data, inner_rtol,
            max_outer, nit.ctypes.data, ini.ctypes.data, fin.ctypes.data, C.byref(outer), _stream(stream)))
        return dict(niter=nit, ini_res=ini, fin_res=fin, outer=outer.value)

    def split(self, v):
        """Per-system views of a batch vector."""
        return [v[self.offsets[k]:self.offsets[k + 1]] for k in range(self.k)]

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.tflows_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


CALCPS_FIELDS = ("asave", "rho_n", "rho_o", "rho_oo", "u", "v", "w", "p", "px", "py", "pz")


class UvwParams(C.Structure):
    """tflows_uvw_params: the flags and constants of one newuvw call."""
    _fields_ = [("var", C.c_int32), ("convec", C.c_int32), ("diffus", C.c_int32), ("cross", C.c_int32),
                ("inert", C.c_int32), ("visc", C.c_double), ("dt", C.c_double), ("urf", C.c_double),
                ("blend", C.c_int32 * 8), ("blend_tem", C.c_int32 * 8), ("urfc", C.c_double * 8),
                ("visc_rho", C.c_int32)]

    @classmethod
    def from_dict(cls, p):
        q = cls()
        for k in ("var", "convec", "diffus", "cross", "inert"):
            setattr(q, k, int(p[k]))
        for k in ("visc", "dt", "urf"):
            setattr(q, k, float(p[k]))
        q.visc_rho = int(p.get("visc_rho", 0))
        if len(p["blend"]) > 8:
            raise ValueError("at most 8 materials")
        for m in range(len(p["blend"])):
            q.blend[m] = int(p["blend"][m])
            q.blend_tem[m] = int(p["blend_tem"][m])
            q.urfc[m] = float(p["urfc"][m])
        return q


class BFile:
    """NAME.b, the boundary-condition file, read by the library
    (csrc/meshio.cpp, after Process/BouLoa.f90:30-160).  ``hot`` is HOT =
    YES.  Attributes: nmat, visc (the last material's VISc), statemat (FLUID
    7, SOLID 8), denc, conc, capc; nb, typbou (allp_mod codes, file order),
    bfile (1 where the condition is a FILE profile), profiles (names, '' for
    constants), bu, bv, bw, bp, bz; ninit, ifile, inits (names), iu, iv, iw,
    iz.  Host only."""

    def __init__(self, path, hot=False):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.tflows_bfile_load(os.fsencode(path), 1 if hot else 0, C.byref(h)))
        try:
            n = [C.c_int32(0) for _ in range(3)]
            visc = C.c_double(0.0)
            _check(lib.tflows_bfile_info(h, *(C.byref(x) for x in n), C.byref(visc)))
            self.nmat, self.nb, self.ninit = (x.value for x in n)
            self.visc = visc.value
            i32 = lambda k: np.zeros(max(k, 1), np.int32)        # noqa: E731
            f64 = lambda k: np.zeros(max(k, 1), np.float64)      # noqa: E731
            self.statemat, self.denc, self.conc, self.capc = i32(self.nmat), f64(self.nmat), f64(self.nmat), \
                f64(self.nmat)
            _check(lib.tflows_bfile_materials(h, *(a.ctypes.data for a in (self.statemat, self.denc, self.conc,
                                                                             self.capc))))
            b = [i32(self.nb), i32(self.nb)] + [f64(self.nb) for _ in range(5)]
            _check(lib.tflows_bfile_bounds(h, *(a.ctypes.data for a in b)))
            self.typbou, self.bfile, self.bu, self.bv, self.bw, self.bp, self.bz = (a[:self.nb] for a in b)
            q = [i32(self.ninit)] + [f64(self.ninit) for _ in range(4)]
            _check(lib.tflows_bfile_inits(h, *(a.ctypes.data for a in q)))
            self.ifile, self.iu, self.iv, self.iw, self.iz = (a[:self.ninit] for a in q)
            for k in ("statemat", "denc", "conc", "capc"):
                setattr(self, k, getattr(self, k)[:self.nmat])
            buf = C.create_string_buffer(512)

            def name(kind, k):
                _check(lib.tflows_bfile_name(h, kind, k, buf, len(buf)))
                return buf.value.decode()
            self.profiles = [name(0, k) for k in range(1, self.nb + 1)]
            self.inits = [name(1, k) for k in range(1, self.ninit + 1)]
        finally:
            lib.tflows_bfile_destroy(h)


class Mesh:
    """A T-FlowS mesh read from NAME.cns and NAME.geo by the library's reader
    (csrc/meshio.cpp: Process/CnsLoa.f90, Library/GeoLoa.f90, with the
    pattern of Process/TopolM.f90 and the centre distances of
    Process/Calc3.f90).  Host only.  Arrays are numpy copies in the
    reference's layout: cell arrays (-NbC:NC) at [NbC + c], ``sidec`` as
    (3, NS) (SideC(0:2,NS)), ``sidaij`` as (2, NS), 1-based indices."""

    INT = ("material", "bcmark", "copyc", "sidec", "acol", "arow", "adia", "sidaij")
    REAL = ("xc", "yc", "zc", "volume", "delta", "wallds", "sx", "sy", "sz", "f", "xsp", "ysp", "zsp",
            "dx_file", "dy_file", "dz_file", "dx", "dy", "dz", "scoef")

    def __init__(self, cns_path, geo_path=None):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.tflows_mesh_load(os.fsencode(cns_path), os.fsencode(geo_path) if geo_path else None,
                                         C.byref(h)))
        self.handle = h
        v = [C.c_int32(0) for _ in range(6)]
        _check(self.lib.tflows_mesh_info(h, *(C.byref(x) for x in v)))
        self.nc, self.nbc, self.ns, self.nmat, self.nonz, geo = (x.value for x in v)
        self.has_geo = bool(geo)
        ncell = self.nbc + self.nc + 1
        sizes = {"material": ncell, "bcmark": ncell, "copyc": ncell, "sidec": 3 * self.ns, "acol": self.nc + 1,
                 "arow": self.nonz, "adia": self.nc, "sidaij": 2 * self.ns}
        for k in self.INT:
            a = np.empty(sizes[k], np.int32)
            _check(self.lib.tflows_mesh_int_array(h, k.encode(), a.ctypes.data, a.size))
            setattr(self, k, a)
        self.sidec = self.sidec.reshape(self.ns, 3).T.copy()
        self.sidaij = self.sidaij.reshape(self.ns, 2).T.copy()
        if self.has_geo:
            for k in self.REAL:
                n = ncell if k in ("xc", "yc", "zc", "volume", "delta", "wallds") else self.ns
                a = np.empty(n, np.float64)
                _check(self.lib.tflows_mesh_real_array(h, k.encode(), a.ctypes.data, a.size))
                setattr(self, k, a)

    def typebc(self, typbou):
        """TypeBC(-NbC:NC) for the boundary types typBou(1:NB) of the .b
        file (codes of meshgen.BC_CODES)."""
        t = np.ascontiguousarray(typbou, np.int32)
        out = np.empty(self.nbc + self.nc + 1, np.int32)
        _check(self.lib.tflows_mesh_typebc(self.handle, t.ctypes.data, len(t), out.ctypes.data, out.size))
        return out

    def ff(self, typebc, statemat=None):
        """fF(NS) (Calc3.f90:53-71); statemat: StateMat(1:Nmat), FLUID 7 /
        SOLID 8 (None: all fluid)."""
        t = np.ascontiguousarray(typebc, np.int32)
        st = None if statemat is None else np.ascontiguousarray(statemat, np.int32)
        out = np.empty(self.ns, np.float64)
        _check(self.lib.tflows_mesh_ff(self.handle, t.ctypes.data, None if st is None else st.ctypes.data,
                                       out.ctypes.data, out.size))
        return out

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.tflows_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceCalcPS:
    """The pressure-correction assembly of one mesh (Process/CalcPS.f90:51-202)
    on the GPU.  ``case`` is a meshgen.CalcPSCase (reference layout); the
    face list and geometry go to the device once."""

    def __init__(self, case, stream=None):
        import torch
        self.lib = load_library()
        self.case = case
        dev = torch.device("cuda")
        t32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)    # noqa: E731
        t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)  # noqa: E731
        keep = [t32(case.sidec.T), t32(case.sidaij.T), t32(case.adia), t32(case.typebc)]
        keep += [t64(getattr(case, k)) for k in ("sx", "sy", "sz", "dx", "dy", "dz", "f", "volume")]
        torch.cuda.synchronize()
        h = C.c_void_p()
        _check(self.lib.tflows_calcps_create(case.nc, case.nbc, case.ns, case.nonz, *(_ptr(t) for t in keep), 1,
                                             _stream(stream), C.byref(h)))
        self.handle = h

    def visit_order(self) -> int:
        """How the momentum and scalar cell kernels visit the cells: 2 blocks
        of 64 cells x 2 lines x 2 planes, 1 tiles brick by brick, 0 tiles in
        order."""
        v = int(self.lib.tflows_calcps_visit_order(self.handle))
        if v < 0:
            _check(v)
        return v

    def assemble(self, fields=None, nmat=None, dt=None, out=None, stream=None):
        """Assemble from cuda float64 field tensors ((-NbC:NC) layout; the
        case's own fields by default).  Returns dict of cuda tensors aval,
        b, flux, abou and floats errmax, mean; ``out`` reuses tensors."""
        import torch
        c = self.case
        if fields is None:
            fields = {k: torch.from_numpy(getattr(c, k)).cuda() for k in CALCPS_FIELDS}
        if out is None:
            dev = torch.device("cuda")
            out = dict(aval=torch.empty(c.nonz, dtype=torch.float64, device=dev),
                       b=torch.empty(c.nc, dtype=torch.float64, device=dev),
                       flux=torch.zeros(max(c.ns, 1), dtype=torch.float64, device=dev),
                       abou=torch.zeros(max(c.nbc, 1), dtype=torch.float64, device=dev))
        em, mn = C.c_double(0.0), C.c_double(0.0)
        _check(self.lib.tflows_calcps_assemble(
            self.handle, c.nmat if nmat is None else nmat, c.dt if dt is None else dt,
            *(_ptr(fields[k]) for k in CALCPS_FIELDS), _ptr(out["aval"]), _ptr(out["b"]), _ptr(out["flux"]),
            _ptr(out["abou"]), C.byref(em), C.byref(mn), 1, _stream(stream)))
        return dict(out, errmax=em.value, mean=mn.value)

    def assemble_host(self, nmat=None, dt=None):
        """The same from the case's host arrays (copied in and out)."""
        c = self.case
        aval, b = np.empty(c.nonz), np.empty(c.nc)
        flux, abou = np.zeros(max(c.ns, 1)), np.zeros(max(c.nbc, 1))
        em, mn = C.c_double(0.0), C.c_double(0.0)
        fl = [np.ascontiguousarray(getattr(c, k), np.float64) for k in CALCPS_FIELDS]
        _check(self.lib.tflows_calcps_assemble(
            self.handle, c.nmat if nmat is None else nmat, c.dt if dt is None else dt,
            *(a.ctypes.data for a in fl), aval.ctypes.data, b.ctypes.data, flux.ctypes.data, abou.ctypes.data,
            C.byref(em), C.byref(mn), 0, None))
        return dict(aval=aval, b=b, flux=flux[:c.ns], abou=abou[:c.nbc], errmax=em.value, mean=mn.value)

    def correct(self, pp, aval, abou, u, v, w, flux, fields=None, dt=None, b=None, stream=None):
        """The correction after the pressure solve (CorUVW.f90:36-121) on
        cuda float64 tensors: u, v, w ((-NbC:NC)) and flux (NS) in place;
        returns (b, errmax), b the mass error per unit volume (NC)."""
        import torch
        c = self.case
        if fields is None:
            fields = {k: torch.from_numpy(getattr(c, k)).cuda() for k in CALCPS_FIELDS}
        if b is None:
            b = torch.empty(c.nc, dtype=torch.float64, device="cuda")
        em = C.c_double(0.0)
        _check(self.lib.tflows_calcps_correct(
            self.handle, c.dt if dt is None else dt, _ptr(pp), _ptr(aval), _ptr(abou),
            *(_ptr(fields[k]) for k in ("asave", "rho_n", "rho_o", "rho_oo", "px", "py", "pz")),
            _ptr(u), _ptr(v), _ptr(w), _ptr(flux), _ptr(b), C.byref(em), 1, _stream(stream)))
        return b, em.value

    # -- momentum assembly (Process/NewUVW.f90:137-499) --------------------
    def uvw_setup(self, uf):
        """Scoef, fF and the cell materials of meshgen.uvw_fields, once."""
        import torch
        t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()   # noqa: E731
        keep = [t64(uf["scoef"]), t64(uf["ff"]),
                torch.from_numpy(np.ascontiguousarray(uf["material"], np.int32)).cuda()]
        torch.cuda.synchronize()
        _check(self.lib.tflows_calcps_uvw_setup(self.handle, _ptr(keep[0]), _ptr(keep[1]), _ptr(keep[2]),
                                                int(uf["nmaterials"]), 1, None))
        self.nmaterials = int(uf["nmaterials"])

    @staticmethod
    def uvw_inputs(args):
        """cuda float64 tensors of meshgen.uvw_args(...) in the order of
        tflows_calcps_uvw_assemble: 15 cell arrays, Flux_u, then the mass
        flux Flux (None when args has none: Flux_u decides)."""
        import torch
        t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()   # noqa: E731
        cells = [args[k] for k in ("rho_n", "rho_o", "rho_oo", "ui_n", "ui_o", "ui_oo", "source", "pi")]
        fl = args.get("flux")
        return [t64(a) for a in cells + list(args["g"])] + [t64(args["flux_u"]), None if fl is None else t64(fl)]

    def assemble_uvw(self, params, inputs, out=None, stream=None):
        """One component's system from cuda tensors ``inputs`` (uvw_inputs);
        ``params`` a meshgen.uvw_params dict.  Returns dict of cuda tensors
        aval, b, abou, asave ((-NbC:NC), cells 1..NC written for var = 1),
        rhoui ((-NbC:NC))."""
        import torch
        c = self.case
        if out is None:
            dev = torch.device("cuda")
            ncell = c.nbc + c.nc + 1
            out = dict(aval=torch.empty(c.nonz, dtype=torch.float64, device=dev),
                       b=torch.empty(c.nc, dtype=torch.float64, device=dev),
                       abou=torch.zeros(max(c.nbc, 1), dtype=torch.float64, device=dev),
                       asave=torch.zeros(ncell, dtype=torch.float64, device=dev),
                       rhoui=torch.empty(ncell, dtype=torch.float64, device=dev))
        P = UvwParams.from_dict(params)
        _check(self.lib.tflows_calcps_uvw_assemble(
            self.handle, C.byref(P), *(None if t is None else _ptr(t) for t in inputs), _ptr(out["aval"]),
            _ptr(out["b"]),
            _ptr(out["abou"]), _ptr(out["asave"]), _ptr(out["rhoui"]), 1, _stream(stream)))
        return out

    def assemble_uvw_host(self, params, args):
        """The same from host arrays (meshgen.uvw_args), copied in and out."""
        c = self.case
        ncell = c.nbc + c.nc + 1
        cells = [np.ascontiguousarray(args[k], np.float64)
                 for k in ("rho_n", "rho_o", "rho_oo", "ui_n", "ui_o", "ui_oo", "source", "pi")]
        cells += [np.ascontiguousarray(a, np.float64) for a in args["g"]]
        fl = np.ascontiguousarray(args["flux_u"], np.float64)
        mf = None if args.get("flux") is None else np.ascontiguousarray(args["flux"], np.float64)
        aval, b = np.empty(c.nonz), np.empty(c.nc)
        abou, asave, rhoui = np.zeros(max(c.nbc, 1)), np.zeros(ncell), np.empty(ncell)
        P = UvwParams.from_dict(params)
        _check(self.lib.tflows_calcps_uvw_assemble(
            self.handle, C.byref(P), *(a.ctypes.data for a in cells), fl.ctypes.data,
            None if mf is None else mf.ctypes.data, aval.ctypes.data,
            b.ctypes.data, abou.ctypes.data, asave.ctypes.data, rhoui.ctypes.data, 0, None))
        return dict(aval=aval, b=b, abou=abou[:c.nbc], asave=asave, rhoui=rhoui)

    # -- mixture-fraction assembly (Process/CalcZmix.f90:100-387) --------
    @staticmethod
    def zmix_inputs(args):
        """cuda float64 tensors of meshgen.zmix_args(...) in the order of
        tflows_calcps_zmix_assemble: 10 cell arrays, Flux_u, then the mass
        flux Flux (None when args has none)."""
        import torch
        t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()   # noqa: E731
        cells = [args[k] for k in ("rho_n", "rho_o", "rho_oo", "z_n", "z_o", "z_oo", "source")]
        fl = args.get("flux")
        return [t64(a) for a in cells + list(args["g"])] + [t64(args["flux_u"]), None if fl is None else t64(fl)]

    def assemble_zmix(self, params, inputs, out=None, stream=None):
        """The mixture-fraction system from cuda tensors ``inputs``
        (zmix_inputs); returns dict of cuda tensors aval, b, abou, rhoz."""
        import torch
        c = self.case
        if out is None:
            dev = torch.device("cuda")
            out = dict(aval=torch.empty(c.nonz, dtype=torch.float64, device=dev),
                       b=torch.empty(c.nc, dtype=torch.float64, device=dev),
                       abou=torch.zeros(max(c.nbc, 1), dtype=torch.float64, device=dev),
                       rhoz=torch.empty(c.nbc + c.nc + 1, dtype=torch.float64, device=dev))
        P = UvwParams.from_dict(params)
        _check(self.lib.tflows_calcps_zmix_assemble(
            self.handle, C.byref(P), *(None if t is None else _ptr(t) for t in inputs), _ptr(out["aval"]),
            _ptr(out["b"]),
            _ptr(out["abou"]), _ptr(out["rhoz"]), 1, _stream(stream)))
        return out

    def assemble_zmix_host(self, params, args):
        """The same from host arrays (meshgen.zmix_args), copied in and out."""
        c = self.case
        cells = [np.ascontiguousarray(args[k], np.float64)
                 for k in ("rho_n", "rho_o", "rho_oo", "z_n", "z_o", "z_oo", "source")]
        cells += [np.ascontiguousarray(a, np.float64) for a in args["g"]]
        fl = np.ascontiguousarray(args["flux_u"], np.float64)
        mf = None if args.get("flux") is None else np.ascontiguousarray(args["flux"], np.float64)
        aval, b = np.empty(c.nonz), np.empty(c.nc)
        abou, rhoz = np.zeros(max(c.nbc, 1)), np.empty(c.nbc + c.nc + 1)
        P = UvwParams.from_dict(params)
        _check(self.lib.tflows_calcps_zmix_assemble(
            self.handle, C.byref(P), *(a.ctypes.data for a in cells), fl.ctypes.data,
            None if mf is None else mf.ctypes.data, aval.ctypes.data,
            b.ctypes.data, abou.ctypes.data, rhoz.ctypes.data, 0, None))
        return dict(aval=aval, b=b, abou=abou[:c.nbc], rhoz=rhoz)

    def cfl_peclet(self, rho_n, flux, dt, visc, stream=None):
        """The CFL and Peclet numbers per material (CorUVW.f90:124-147) from
        cuda float64 rho_n ((-NbC:NC)) and flux (NS), after uvw_setup.
        Returns two numpy arrays of Nmat values."""
        n = getattr(self, "nmaterials", 0)
        cfl, pe = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        _check(self.lib.tflows_calcps_cfl_peclet(self.handle, float(dt), float(visc), _ptr(rho_n), _ptr(flux),
                                                 cfl.ctypes.data, pe.ctypes.data, 1, _stream(stream)))
        return cfl[:n], pe[:n]

    def zmix_update(self, rhoz, rho_n, z_n, stream=None):
        """Zmix%n(1:NC) = rhoZ / rho%n limited to [0, 1] on cuda tensors."""
        _check(self.lib.tflows_calcps_zmix_update(self.handle, _ptr(rhoz), _ptr(rho_n), _ptr(z_n), 1,
                                                  _stream(stream)))

    def uvw_update(self, rhoui, rho_n, ui_n, stream=None):
        """Ui%n(1:NC) = rhoUi / rho%n on cuda tensors (NewUVW.f90:526)."""
        _check(self.lib.tflows_calcps_uvw_update(self.handle, _ptr(rhoui), _ptr(rho_n), _ptr(ui_n), 1,
                                                 _stream(stream)))

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.tflows_calcps_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
