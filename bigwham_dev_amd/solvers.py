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
                 correction solves H d = r by one restart cycle of gmres()
                 driven by the FP32-factor matvec (FP64 Arnoldi and vectors),
                 so every restart of the Krylov cycle starts from a true FP64
                 residual.  Stops on the FP64 residual.
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
          callback=None):
    """Solve A x = b; matvec(v) -> A v on device (float64).  precond(v) -> M^{-1} v
    (right preconditioning).  Returns (x, iterations, rel_residual_estimate)."""
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
    while its < maxiter:
        r = b - matvec(x) if its > 0 or x0 is not None else b.clone()
        beta = torch.linalg.vector_norm(r).item()
        rel = beta / bnorm
        if rel <= tol:
            break
        V[0] = r / beta
        Hs = [[0.0] * restart for _ in range(restart + 1)]
        cs, sn = [0.0] * restart, [0.0] * restart
        g = [0.0] * (restart + 1)
        g[0] = beta
        k_done = 0
        for k in range(restart):
            w = matvec(M(V[k]))
            its += 1
            # CGS2: two classical Gram-Schmidt passes as matrix-vector products
            h = V[: k + 1] @ w
            w = w - V[: k + 1].T @ h
            h2 = V[: k + 1] @ w
            w = w - V[: k + 1].T @ h2
            h = (h + h2).tolist()
            hn = torch.linalg.vector_norm(w).item()
            for i in range(k + 1):
                Hs[i][k] = h[i]
            Hs[k + 1][k] = hn
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
            if hn != 0.0:
                V[k + 1] = w / hn
            if rel <= tol or its >= maxiter or hn == 0.0:
                break
        # back substitution
        y = [0.0] * k_done
        for i in range(k_done - 1, -1, -1):
            s = g[i] - sum(Hs[i][j] * y[j] for j in range(i + 1, k_done))
            y[i] = s / Hs[i][i]
        yt = torch.tensor(y, dtype=dt, device=dev)
        x = x + M(V[:k_done].T @ yt)
        x0 = x  # next restart recomputes the true residual
        if rel <= tol:
            break
    return x, its, rel


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

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, its, _ = gmres(mv, b, M, tol=tol, restart=restart, maxiter=maxiter)
    r = b - hmat.matvec(x, original=True, precision="fp64")
    rel = (torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b)).item()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return x, SolveInfo(rel <= tol, its, n64[0] + 1, 0, rel, dt)


def solve_mixed(hmat, b: torch.Tensor, tol: float = 1e-8, inner_tol: float = 1e-5,
                restart: int = 30, maxiter: int = 2000, max_refine: int = 500,
                precondition: bool = True) -> tuple[torch.Tensor, SolveInfo]:
    """Iterative refinement: FP64 residuals, FP32-factor inner GMRES.

    Each refinement step runs one restart cycle (at most `restart` FP32
    matvecs) on H d = r, stopping early once the inner residual drops below
    max(inner_tol, 0.5 tol / ||r||/||b||); the next step starts from the
    exact FP64 residual.  The iteration count then follows the FP64 restarted
    GMRES, with one extra FP64 matvec per cycle."""
    d = torch.from_numpy(hmat.diagonal(original=True)).to(b.device)
    M = jacobi(d) if precondition else None
    counts = {"fp64": 0, "fp32": 0}

    def mv32(v):
        counts["fp32"] += 1
        return hmat.matvec(v, original=True, precision="fp32")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bnorm = torch.linalg.vector_norm(b).item()
    x = torch.zeros_like(b)
    r = b.clone()
    rel = 1.0
    hist = []
    its = 0
    for _ in range(max_refine):
        target = max(inner_tol, 0.5 * tol / rel)
        dx, k, _ = gmres(mv32, r, M, tol=target, restart=restart,
                         maxiter=min(restart, maxiter - its))
        its += k
        x = x + dx
        r = b - hmat.matvec(x, original=True, precision="fp64")
        counts["fp64"] += 1
        rel = torch.linalg.vector_norm(r).item() / bnorm
        hist.append(rel)
        if rel <= tol or its >= maxiter:
            break
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return x, SolveInfo(rel <= tol, its, counts["fp64"], counts["fp32"], rel, dt, hist)
