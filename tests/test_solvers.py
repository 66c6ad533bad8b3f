"""CPU checks of the device solvers (bigwham_dev_amd.solvers) on dense
operators: restarted GMRES against a direct solve, the Arnoldi-relation
restart against the explicit residual, and the mixed-precision refinement
with an FP32 operator reaching an FP64-verified residual.  The reference has
no solver of its own (users call scipy gmres on BEMatrix with the Jacobi
preconditioner, interfaces/python/bigwham4py/bigwham4py.py:362-380)."""
import numpy as np
import pytest
import torch

from bigwham_dev_amd import solvers


def _system(n=300, seed=0):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 1.3 + rng.standard_normal((n, n)) / np.sqrt(n)
    A += np.diag(rng.uniform(0.5, 3.0, n))  # uneven diagonal: Jacobi helps
    b = rng.standard_normal(n)
    return torch.from_numpy(A), torch.from_numpy(b)


@pytest.mark.parametrize("implicit", [False, True])
@pytest.mark.parametrize("restart", [5, 20])
def test_gmres_dense(implicit, restart):
    A, b = _system()
    M = solvers.jacobi(torch.diagonal(A).clone())
    x, its, rel = solvers.gmres(lambda v: A @ v, b, M, tol=1e-10, restart=restart,
                                maxiter=500, implicit_restart=implicit)
    true_rel = (torch.linalg.vector_norm(b - A @ x) / torch.linalg.vector_norm(b)).item()
    assert rel <= 1e-10
    assert true_rel <= 5e-10
    x_ref = np.linalg.solve(A.numpy(), b.numpy())
    assert np.linalg.norm(x.numpy() - x_ref) / np.linalg.norm(x_ref) < 1e-8
    assert its > restart  # several restart cycles were exercised


def test_implicit_restart_residual_matches_explicit():
    """After one cycle the Arnoldi residual V(beta e1 - Hbar y) is b - A x."""
    A, b = _system(200, 1)
    seen = []

    def mv(v):
        seen.append(v.clone())
        return A @ v

    x, its, _ = solvers.gmres(mv, b, None, tol=0.0, restart=10, maxiter=11,
                              implicit_restart=True)
    # the 11th matvec was applied to V[0] of cycle 2 = r_implicit / ||r_implicit||
    v0 = seen[10]
    r = b - A @ solvers.gmres(lambda v: A @ v, b, None, tol=0.0, restart=10, maxiter=10)[0]
    r_dir = r / torch.linalg.vector_norm(r)
    assert torch.allclose(v0, r_dir, atol=1e-12)


def test_gmres_zero_rhs_and_exact_breakdown():
    A, _ = _system(50, 2)
    x, its, rel = solvers.gmres(lambda v: A @ v, torch.zeros(50, dtype=torch.float64))
    assert its == 0 and float(x.abs().max()) == 0.0
    # identity operator: Krylov space of dimension 1, breakdown after one step
    b = torch.arange(1.0, 51.0, dtype=torch.float64)
    x, its, rel = solvers.gmres(lambda v: v.clone(), b, None, tol=1e-14, restart=10,
                                implicit_restart=True)
    assert its == 1 and torch.allclose(x, b, rtol=0, atol=1e-12)


class _DenseOperator:
    """Stand-in for HipHmat on the CPU: FP64 and FP32 products of one matrix."""

    def __init__(self, A):
        self.A = A
        self.A32 = A.float()
        self.precisions = ("fp64", "fp32")

    def matvec(self, v, original=True, precision="fp64"):
        if precision == "fp32":
            return (self.A32 @ v.float()).double()
        return self.A @ v

    def diagonal(self, original=True):
        return torch.diagonal(self.A).numpy().copy()


def test_mixed_refinement_reaches_fp64_tolerance():
    A, b = _system(400, 3)
    op = _DenseOperator(A)
    x, info = solvers.solve_mixed(op, b, tol=1e-10, restart=20)
    assert info.converged and info.rel_residual <= 1e-10
    true_rel = (torch.linalg.vector_norm(b - A @ x) / torch.linalg.vector_norm(b)).item()
    assert true_rel == pytest.approx(info.rel_residual, rel=1e-6)
    assert 2 <= info.matvecs_fp64 <= 4          # FP64 residuals only between refinements
    assert info.matvecs_fp32 == info.iterations
    # the FP32 operator alone cannot get there: a plain FP32 solve stalls near 1e-7
    x32, _, _ = solvers.gmres(lambda v: op.matvec(v, precision="fp32"), b,
                              solvers.jacobi(torch.diagonal(A).clone()), tol=1e-12,
                              restart=20, maxiter=300)
    r32 = (torch.linalg.vector_norm(b - A @ x32) / torch.linalg.vector_norm(b)).item()
    assert r32 > 1e-9
    x64, info64 = solvers.solve_fp64(op, b, tol=1e-10, restart=20)
    assert info64.converged
    assert np.linalg.norm((x - x64).numpy()) / np.linalg.norm(x64.numpy()) < 1e-8
