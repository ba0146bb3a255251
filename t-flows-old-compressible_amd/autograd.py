"""Differentiable Krylov solve for gradient checks (PyTorch-ROCm).

:class:`KrylovSolve` is a thin ``torch.autograd.Function`` over the HIP
library: the forward pass solves A x = b with the reference's CG / BiCG /
CGS on the GPU; the backward pass solves the adjoint system A^T lambda =
dL/dx with the same method (tflows_solve_adjoint: the values of A^T are
gathered from the mirror positions of the structurally symmetric pattern)
and forms

    dL/db = lambda,      dL/dA_ij = -lambda_i x_j   on the stored pattern

with the outer-product kernel (tflows_outer_grad).  It exists for the
gradient harness; the solver itself needs no torch.
"""
from __future__ import annotations

import torch


class KrylovSolve(torch.autograd.Function):
    """x = solve(aval, b) for a DeviceSystem's pattern.

    aval: cuda float64 tensor of length NONZ (the T-FlowS value array);
    b: cuda float64 tensor of length N.  The solve runs to
    ||b - A x|| < rtol ||b|| (the reference's absolute test with
    tol = rtol ||b||); non-convergence raises.
    """

    @staticmethod
    def forward(ctx, aval, b, system, method="cg", prec=1, rtol=1e-12, cap=100000):
        system.set_values(aval)
        x = torch.zeros_like(b)
        r = b.clone()
        tol = rtol * float(torch.linalg.vector_norm(b))
        it, ini, fin = system.solve(method, prec, x, r, cap, tol)
        if not fin < tol:
            raise RuntimeError(f"forward solve did not converge: FinRes={fin:.3e} tol={tol:.3e} niter={it}")
        ctx.system, ctx.method, ctx.prec, ctx.rtol, ctx.cap = system, method, prec, rtol, cap
        ctx.save_for_backward(aval, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        aval, x = ctx.saved_tensors
        system = ctx.system
        system.set_values(aval)
        lam = torch.zeros_like(x)
        r = gx.contiguous().clone()
        gnorm = float(torch.linalg.vector_norm(r))
        if gnorm > 0.0:
            tol = ctx.rtol * gnorm
            it, ini, fin = system.solve_adjoint(ctx.method, ctx.prec, lam, r, ctx.cap, tol)
            if not fin < tol:
                raise RuntimeError(f"adjoint solve did not converge: FinRes={fin:.3e} tol={tol:.3e}")
        grad_a = torch.empty_like(aval)
        system.outer_grad(lam, x, grad_a)
        return grad_a, lam, None, None, None, None, None


def solve(aval, b, system, method="cg", prec=1, rtol=1e-12, cap=100000):
    return KrylovSolve.apply(aval, b, system, method, prec, rtol, cap)
