"""Gradient harness of the adjoint solve on a perturbed, renumbered mesh.

Problem: the pressure-correction matrix of an n^3 hex box with tanh wall
stretching, random node perturbation and random cell renumbering, its face
coefficients a_f = |S|^2/(S.d) computed from the perturbed geometry in FP64;
loss L = 0.5 ||x - x_t||^2 with x = A(a)^-1 b and x_t a sum of 4 seeded
Fourier modes.  Gradients with respect to b and to the face coefficients a_f
come from :class:`autograd.KrylovSolve` (adjoint CG + outer-product kernel;
torch's autograd carries dL/dA to dL/da_f through the assembly) and are
compared with

  * central finite differences of L (step h on each checked entry), and
  * an FP64 CPU reference gradient: x and lambda = A^-T (x - x_t) solved on
    the host (scipy's direct solver up to 60,000 rows, beyond it a Jacobi
    PCG on torch's CPU sparse product), then dL/db = lambda and
    dL/da_f = -(lambda_c1 - lambda_c2)(x_c1 - x_c2).

Finite-difference criteria.  Over the checked entries as a vector:
||ad - fd|| <= 1e-6 ||fd||.  Per entry a central difference carries the
error of the two losses it subtracts.  Each loss comes from a solve that
stops at ||r|| < rtol ||b||, which moves it by lambda . r (lambda the
adjoint state), at most ||lambda|| rtol ||b||, plus the rounding of L,
eps L; so |fd - dL/dp| <= (||lambda|| rtol ||b|| + eps L) / h, beside the
truncation term (h^2 / 6) d3L/dp3, which is zero for the entries of b (L
is quadratic in b) and was measured negligible for the coefficients.  Per
entry: |ad - fd| <= 1e-6 |fd| + that bound (fd_floor_ratio is the largest
ratio of the two sides); the cancellation floor eps L / h is reported
beside it.
"""
from __future__ import annotations

import time

import numpy as np


def build_problem(pkg, n: int, seed: int = 4, ratio: float = 1.05, perturb: float = 0.2):
    mg = pkg.meshgen
    geo = mg.perturbed_box_geometry(n, ratio=ratio, perturb=perturb, seed=seed)
    geo, perm = mg.renumber_faces(geo, seed)
    topo = mg.topolm(n ** 3, geo.c1, geo.c2)
    aval = mg.assemble_face_values(topo, geo.c1, geo.c2, geo.coef, 1, geo.pin_coef)
    m = mg.TFlowSMatrix(n ** 3, topo.nonz, topo.acol, topo.arow, topo.adia, aval)
    b = mg.seeded_rhs(m.n, seed, True)
    xt = mg.fourier_target(geo.centres, seed + 100)
    return dict(geo=geo, topo=topo, matrix=m, b=b, xt=xt)


def torch_assembly(torch, prob, device):
    """Differentiable map a_f -> aval (index_add on the SidAij / Adia slots)."""
    topo, geo = prob["topo"], prob["geo"]
    t = lambda a, dt=torch.int64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
    s1, s2 = t(topo.sidaij[0]), t(topo.sidaij[1])
    d1, d2 = t(topo.adia[geo.c1 - 1] - 1), t(topo.adia[geo.c2 - 1] - 1)
    pin_pos = int(topo.adia[0] - 1)
    nonz = topo.nonz

    def assemble(coef):
        aval = torch.zeros(nonz, dtype=torch.float64, device=device)
        aval = aval.index_add(0, s1, -coef).index_add(0, s2, -coef)
        aval = aval.index_add(0, d1, coef).index_add(0, d2, coef)
        pin = torch.zeros_like(aval)
        pin[pin_pos] = geo.pin_coef
        return aval + pin
    return assemble


def _cpu_pcg(A, b, rtol, maxit=200000, log=None):
    """Jacobi-preconditioned CG in FP64 on the host, with torch's
    multithreaded CPU sparse product (scipy's is single-threaded and takes
    minutes at 96^3); stops when the recurrence residual reaches
    ||r|| <= rtol ||b||, as CG.f90:248-255 and scipy's cg do."""
    import torch
    At = torch.sparse_csr_tensor(torch.from_numpy(A.indptr.astype(np.int64)),
                                 torch.from_numpy(A.indices.astype(np.int64)),
                                 torch.from_numpy(np.ascontiguousarray(A.data)), size=A.shape)
    d = torch.from_numpy(np.ascontiguousarray(A.diagonal()))
    bt = torch.from_numpy(np.ascontiguousarray(b))
    x = torch.zeros_like(bt)
    r = bt.clone()
    z = r / d
    p = z.clone()
    rz = float(torch.dot(r, z))
    stop = rtol * float(torch.linalg.norm(bt))
    for it in range(maxit):
        if log and it % 1000 == 0:
            log(f"[adjoint] CPU CG iteration {it}: |r| / |b| = {float(torch.linalg.norm(r)) / stop * rtol:.3e}")
        q = At @ p
        a = rz / float(torch.dot(p, q))
        x.add_(p, alpha=a)
        r.sub_(q, alpha=a)
        if float(torch.linalg.norm(r)) <= stop:
            return x.numpy()
        z = r / d
        rz_new = float(torch.dot(r, z))
        p.mul_(rz_new / rz).add_(z)
        rz = rz_new
    raise RuntimeError("CPU reference CG did not converge")


def cpu_reference_gradient(prob, rtol=1e-13, log=None):
    """FP64 CPU gradient: direct solves (scipy) up to 60,000 rows, Jacobi
    PCG on the host beyond."""
    import scipy.sparse.linalg as spla
    A = prob["matrix"].to_scipy().tocsc()
    b, xt, geo = prob["b"], prob["xt"], prob["geo"]
    if A.shape[0] <= 60000:
        x = spla.spsolve(A, b)
        lam = spla.spsolve(A.T.tocsc(), x - xt)
    else:
        x = _cpu_pcg(A.tocsr(), b, rtol, log=log)
        lam = _cpu_pcg(A.T.tocsr(), x - xt, rtol, log=log)
    i, j = geo.c1 - 1, geo.c2 - 1
    ga = -(lam[i] - lam[j]) * (x[i] - x[j])
    return x, lam, ga


def run(pkg, n=16, seed=4, n_check=50, h=1e-6, rtol=1e-13, method="cg", prec=1, reference=True, log=print):
    import torch
    from .autograd import solve

    dev = torch.device("cuda")
    t0 = time.time()
    prob = build_problem(pkg, n, seed)
    m = prob["matrix"]
    ds = pkg.krylov.DeviceSystem(m)
    assemble = torch_assembly(torch, prob, dev)
    coef = torch.tensor(prob["geo"].coef, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.tensor(prob["b"], dtype=torch.float64, device=dev, requires_grad=True)
    xt = torch.tensor(prob["xt"], dtype=torch.float64, device=dev)

    def loss(c, bb):
        x = solve(assemble(c), bb, ds, method, prec, rtol)
        return 0.5 * torch.sum((x - xt) ** 2), x

    L, x = loss(coef, b)
    L.backward()
    g_coef = coef.grad.detach().cpu().numpy()
    g_b = b.grad.detach().cpu().numpy()
    Lv = float(L.detach())
    log(f"[adjoint] n={n} N={m.n} faces={len(g_coef)} L={Lv:.6e} setup+grad {time.time()-t0:.1f}s")

    rng = np.random.default_rng(seed + 7)
    kb = rng.choice(m.n, n_check // 2, replace=False)
    kf = rng.choice(len(g_coef), n_check - n_check // 2, replace=False)
    fd_b, fd_f = [], []
    t1 = time.time()

    def progress(done):
        if done % 10 == 0:
            log(f"[adjoint] central differences {done}/{n_check} ({time.time() - t1:.1f}s)")
    with torch.no_grad():
        for k in kb:
            e = torch.zeros_like(b)
            e[k] = h
            fd_b.append((float(loss(coef, b + e)[0]) - float(loss(coef, b - e)[0])) / (2 * h))
            progress(len(fd_b))
        for k in kf:
            e = torch.zeros_like(coef)
            e[k] = h
            fd_f.append((float(loss(coef + e, b)[0]) - float(loss(coef - e, b)[0])) / (2 * h))
            progress(len(fd_b) + len(fd_f))
    fd = np.concatenate([fd_b, fd_f])
    ad = np.concatenate([g_b[kb], g_coef[kf]])
    scale = np.abs(fd).max()
    rel_fd = np.abs(ad - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    eps = float(np.finfo(float).eps)
    floor_eps = eps * Lv / h                                   # cancellation floor
    floor = (float(np.linalg.norm(g_b)) * rtol * float(np.linalg.norm(prob["b"])) + eps * Lv) / h
    out = dict(n=n, N=m.n, faces=int(len(g_coef)), loss=Lv,
               fd_rel=float(np.linalg.norm(ad - fd) / np.linalg.norm(fd)),
               fd_max_rel=float(rel_fd.max()), fd_median_rel=float(np.median(rel_fd)),
               fd_cancellation_floor=floor_eps, fd_error_bound=floor,
               fd_max_abs_err=float(np.abs(ad - fd).max()),
               fd_max_err_over_cancellation_floor=float(np.abs(ad - fd).max() / floor_eps),
               fd_floor_ratio=float(np.max(np.abs(ad - fd) / (1e-6 * np.abs(fd) + floor))),
               n_checked=int(len(fd)), h=h)
    if reference:
        t2 = time.time()
        log("[adjoint] FP64 CPU reference gradient")
        x_ref, lam_ref, ga_ref = cpu_reference_gradient(prob, log=log)
        log(f"[adjoint] CPU reference gradient done ({time.time() - t2:.1f}s)")
        out["ref_rel_grad_b"] = float(np.linalg.norm(g_b - lam_ref) / np.linalg.norm(lam_ref))
        out["ref_rel_grad_coef"] = float(np.linalg.norm(g_coef - ga_ref) / np.linalg.norm(ga_ref))
        out["ref_rel_x"] = float(np.linalg.norm(x.detach().cpu().numpy() - x_ref) / np.linalg.norm(x_ref))
    log(f"[adjoint] {out}")
    ds.close()
    return out
