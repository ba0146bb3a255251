"""Write the golden fixtures of the Krylov path from the reference solvers.

Run in a container where the reference tree is mounted:

    make -C oracle            # builds oracle/_ref/libtflows_ref.so from the
                              # reference's own Fortran sources
    python tests/golden/make_golden.py

The reference's Process/CG.f90, BiCG.f90, CGS.f90, Prec1.f90 and Resid.f90
(compiled with the flags of Process/Makefile:130) are called through their
Fortran ABI on inputs rebuilt from tests/golden/cases.py.  Written files hold
outputs only (x, final residual r1, niter, IniRes, FinRes, D of Prec1,
r1 - A x of Resid) and the sha256 of the inputs.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

import __graft_entry__  # noqa: E402
from oracle.oracle import Reference  # noqa: E402

import cases  # noqa: E402

SENTINEL = -1.2345e300


def input_digest(m, b, x0):
    from tflows_amd.meshgen import vector_digest
    return vector_digest(m.acol, m.arow, m.adia, m.aval, b, x0 if x0 is not None else np.zeros(0))


def main():
    pkg = __graft_entry__.load()
    ref = Reference()
    out = {}
    for name in cases.solver_cases():
        m, b, x0, nb, prec, niter, tol, method, stale = cases.build(pkg, name)
        if stale is not None:
            ref.set_error(stale)
        x = np.zeros(nb + m.n + 1)
        if x0 is not None:
            x[nb + 1:] = x0
        r1 = b.copy()
        it, ini, fin = ref.solve(method, m.n, nb, m.nonz, m.aval, m.acol, m.arow, m.adia, np.zeros(max(nb, 1)),
                                 x, r1, prec, niter, tol)
        key = name.replace(":", "__").replace("/", "_")
        out[key + "__x"] = x[nb + 1:]
        out[key + "__r"] = r1
        out[key + "__scalars"] = np.array([it, SENTINEL if ini is None else ini, fin, tol])
        out[key + "__digest"] = np.frombuffer(bytes.fromhex(input_digest(m, b, x0)), dtype=np.uint8)
        print(f"{name:28s} niter={it:5d} IniRes={ini} FinRes={fin:.3e}")
    # Prec1 and Resid on two matrices
    for kind, s in (("p10", pkg.meshgen.pressure_system(10, seed=21)),
                    ("m8", pkg.meshgen.momentum_system(8, seed=24))):
        m = s.matrix
        for prec in (0, 1, 2):
            out[f"prec1__{kind}_{prec}"] = ref.prec1(m.n, m.nonz, m.aval, m.acol, m.arow, m.adia, prec)
        xx = np.zeros(m.n + 1)
        xx[1:] = np.random.default_rng(29).standard_normal(m.n)
        r = s.b.copy()
        ref.resid(m.n, 0, m.nonz, m.aval, m.acol, m.arow, np.zeros(1), xx, r)
        out[f"resid__{kind}"] = r
    path = os.path.join(HERE, "krylov_reference.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
