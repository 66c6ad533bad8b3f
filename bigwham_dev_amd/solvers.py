"""Krylov solvers around the MI355X H-matvec (the caller side of the path).

The reference ships no solver: users call scipy.sparse.linalg.gmres on the
BEMatrix LinearOperator with the Jacobi preconditioner 1/diag(H)
(interfaces/python/bigwham4py/bigwham4py.py:362-380), one host<->device round
trip per iteration on the CUDA path.  Here GMRES runs on the device with
torch FP64 vectors:

  gmres()        restarted right-preconditioned GMRES, FP64 Arnoldi with
                 classical Gram-Schmidt and one re-orthogonalisation pass,
                 Givens rotations on the host.
  solve_mixed()  mixed precision: iterative refinement whose residual
                 r = b - H x is evaluated with the FP64 matvec and whose
                 correction solves H d = r with gmres() driven by the
                 FP32-factor matvec (FP64 Arnoldi and vectors),
                 restarting from the Arnoldi residual; stops on the FP64
                 residual.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import torch


@dataclass
class SolveInfo:
    converged: bool
    iterations: int           # total inner matvecs
    matvecs_fp64: int
    matvecs_fp32: int
    rel_residual: float       # ||b - H64 x|| / ||b|| at exit
    seconds: float
    history: list = field(default_factory=list)


def gmres(matvec, b: torch.Tensor, precond=None, x0: torch.Tensor | None = None,
          tol: float = 1e-8, restart: int = 30, maxiter: int = 1000,
          callback=None, implicit_restart: bool = False):
    """Solve A x = b; matvec(v) -> A v on device (float64).  precond(v) -> M^{-1} v
    (right preconditioning).  Returns (x, iterations, rel_residual_estimate).

    Each iteration synchronises with the host once (the Hessenberg column and
    the new basis norm come back together); the basis vector itself is
    normalised on the device, and on the GPU the next matvec is already queued
    when the host waits.  A restart recomputes the residual b - A x with
    one matvec, or, with ``implicit_restart``, forms it from the Arnoldi
    relation r = V_{m+1} (beta e_1 - Hbar y) without a matvec (used by the
    inner solves of the mixed-precision refinement, whose outer loop
    recomputes the true FP64 residual anyway)."""
    dev, dt = b.device, b.dtype
    n = b.numel()
    bnorm = torch.linalg.vector_norm(b).item()
    if bnorm == 0.0:
        return torch.zeros_like(b), 0, 0.0
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    M = precond if precond is not None else (lambda v: v)
    its = 0
    rel = 1.0
    V = torch.empty((restart + 1, n), dtype=dt, device=dev)
    # On the GPU the Hessenberg column comes back through pinned memory behind
    # an event, and the next matvec is queued before the host waits for it, so
    # the device is not left idle while the host applies the rotations.
    pipelined = b.is_cuda
    if pipelined:
        col_host = torch.empty(restart + 2, dtype=dt, pin_memory=True)
        ev = torch.cuda.Event()
    r = b.clone() if x0 is None else b - matvec(x)
    while its < maxiter:
        beta = torch.linalg.vector_norm(r).item()
        rel = beta / bnorm
        if rel <= tol:
            break
        V[0] = r / beta
        Hs = [[0.0] * restart for _ in range(restart + 1)]
        Hraw = [[0.0] * restart for _ in range(restart + 1)]
        cs, sn = [0.0] * restart, [0.0] * restart
        g = [0.0] * (restart + 1)
        g[0] = beta
        k_done = 0
        breakdown = False
        pending = None
        for k in range(restart):
            if pending is not None:
                w, pending = pending, None
            else:
                w = matvec(M(V[k]))
            its += 1
            # CGS2: two classical Gram-Schmidt passes as matrix-vector products
            Vk = V[: k + 1]
            h = Vk @ w
            w = torch.addmv(w, Vk.t(), h, alpha=-1.0)
            h2 = Vk @ w
            w = torch.addmv(w, Vk.t(), h2, alpha=-1.0)
            hn_t = torch.linalg.vector_norm(w)
            torch.div(w, hn_t, out=V[k + 1])
            col_t = torch.cat([h + h2, hn_t.reshape(1)])
            if pipelined:
                col_host[: k + 2].copy_(col_t, non_blocking=True)
                ev.record()
                # look ahead unless the cycle ends here or the last estimate is
                # already near the tolerance (a look-ahead past convergence is
                # one wasted matvec)
                if k + 1 < restart and its < maxiter and rel > 1.5 * tol:
                    pending = matvec(M(V[k + 1]))
                ev.synchronize()
                col = col_host[: k + 2].tolist()
            else:
                col = col_t.tolist()
            hn = col[k + 1]
            for i in range(k + 2):
                Hs[i][k] = Hraw[i][k] = col[i]
            for i in range(k):
                t = cs[i] * Hs[i][k] + sn[i] * Hs[i + 1][k]
                Hs[i + 1][k] = -sn[i] * Hs[i][k] + cs[i] * Hs[i + 1][k]
                Hs[i][k] = t
            den = math.hypot(Hs[k][k], Hs[k + 1][k])
            cs[k] = Hs[k][k] / den if den else 1.0
            sn[k] = Hs[k + 1][k] / den if den else 0.0
            Hs[k][k] = cs[k] * Hs[k][k] + sn[k] * Hs[k + 1][k]
            Hs[k + 1][k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            k_done = k + 1
            rel = abs(g[k + 1]) / bnorm
            if callback is not None:
                callback(its, rel)
            breakdown = hn == 0.0
            if rel <= tol or its >= maxiter or breakdown:
                break
        # back substitution
        y = [0.0] * k_done
        for i in range(k_done - 1, -1, -1):
            s_ = g[i] - sum(Hs[i][j] * y[j] for j in range(i + 1, k_done))
            y[i] = s_ / Hs[i][i]
        yt = torch.tensor(y, dtype=dt, device=dev)
        x = x + M(V[:k_done].T @ yt)
        if rel <= tol or its >= maxiter:
            break
        if implicit_restart:
            m1 = k_done if breakdown else k_done + 1
            c = [(beta if i == 0 else 0.0) - sum(Hraw[i][j] * y[j] for j in range(k_done))
                 for i in range(m1)]
            r = V[:m1].T @ torch.tensor(c, dtype=dt, device=dev)
        else:
            r = b - matvec(x)
    return x, its, rel


def _sync(t: torch.Tensor) -> None:
    if t.is_cuda:
        torch.cuda.synchronize(t.device)


def jacobi(diag: torch.Tensor):
    """Jacobi preconditioner 1/diag(H) (bigwham4py.py:373-380)."""
    inv = 1.0 / diag
    return lambda v: inv * v


def warm_up(hmat, b: torch.Tensor, restart: int = 30) -> None:
    """One short restart cycle per precision, so that library initialisation
    (rocBLAS handles for the Gram-Schmidt products) is not timed."""
    for prec in hmat.precisions:
        gmres(lambda v: hmat.matvec(v, original=True, precision=prec), b, None,
              tol=0.0, restart=restart, maxiter=2)
    torch.cuda.synchronize()


def solve_fp64(hmat, b: torch.Tensor, tol: float = 1e-8, restart: int = 30,
               maxiter: int = 2000, precondition: bool = True) -> tuple[torch.Tensor, SolveInfo]:
    """Plain FP64 GMRES on the FP64 H-matvec (the baseline path)."""
    d = torch.from_numpy(hmat.diagonal(original=True)).to(b.device)
    M = jacobi(d) if precondition else None
    n64 = [0]

    def mv(v):
        n64[0] += 1
        return hmat.matvec(v, original=True, precision="fp64")

    _sync(b)
    t0 = time.perf_counter()
    x, its, _ = gmres(mv, b, M, tol=tol, restart=restart, maxiter=maxiter)
    r = b - hmat.matvec(x, original=True, precision="fp64")
    rel = (torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b)).item()
    _sync(b)
    dt = time.perf_counter() - t0
    return x, SolveInfo(rel <= tol, its, n64[0] + 1, 0, rel, dt)


def solve_mixed(hmat, b: torch.Tensor, tol: float = 1e-8, inner_tol: float = 1e-5,
                restart: int = 30, maxiter: int = 2000, max_refine: int = 500,
                precondition: bool = True) -> tuple[torch.Tensor, SolveInfo]:
    """Iterative refinement: FP64 residuals, FP32-factor inner GMRES.

    Each refinement step solves H d = r with restarted GMRES on the FP32-factor
    matvec (FP64 Arnoldi and vectors; restarts take the residual from the
    Arnoldi relation, no extra matvec) down to max(inner_tol, 0.5 tol /
    (||r||/||b||)) of its right-hand side, then recomputes the exact FP64
    residual r = b - H x.  The FP32 operator is accurate to ~1e-7, so each
    refinement can gain about inner_tol; two to three FP64 matvecs reach
    1e-8."""
    d = torch.from_numpy(hmat.diagonal(original=True)).to(b.device)
    M = jacobi(d) if precondition else None
    counts = {"fp64": 0, "fp32": 0}

    def mv32(v):
        counts["fp32"] += 1
        return hmat.matvec(v, original=True, precision="fp32")

    _sync(b)
    t0 = time.perf_counter()
    bnorm = torch.linalg.vector_norm(b).item()
    x = torch.zeros_like(b)
    r = b.clone()
    rel = 1.0
    hist = []
    its = 0
    for _ in range(max_refine):
        target = max(inner_tol, 0.5 * tol / rel)
        dx, k, _ = gmres(mv32, r, M, tol=target, restart=restart, maxiter=maxiter - its,
                         implicit_restart=True)
        its += k
        x = x + dx
        r = b - hmat.matvec(x, original=True, precision="fp64")
        counts["fp64"] += 1
        rel = torch.linalg.vector_norm(r).item() / bnorm
        hist.append(rel)
        if rel <= tol or its >= maxiter:
            break
    _sync(b)
    dt = time.perf_counter() - t0
    return x, SolveInfo(rel <= tol, its, counts["fp64"], counts["fp32"], rel, dt, hist)
