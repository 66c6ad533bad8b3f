"""Adjoint kernel behind torch.autograd: gradients of ||H w - t||^2 against
central finite differences (20 random directions) and against the oracle's
H^T of the same blocks."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HipHmat, HostHmat
from bigwham_dev_amd.meshes import rect_grid

pytestmark = pytest.mark.gpu


def test_gradient_vs_fd_and_dense_adjoint(gpu):
    import torch
    from bigwham_dev_amd.autograd import hmatvec
    coor, conn = rect_grid(40)
    h = HostHmat.from_mesh("3DR0-H-mode1", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    dev = HipHmat(h, precisions=("fp64",), adjoint=True)
    xc = h.collocation_points()
    r2 = xc[:, 0] ** 2 + xc[:, 1] ** 2
    w_true = np.exp(-r2 / 0.2)  # Gaussian-bump opening
    rng = np.random.default_rng(0)
    t_obs = O.matvec_original(h.flat(), w_true)
    t_obs = t_obs + 0.01 * np.abs(t_obs).mean() * rng.standard_normal(t_obs.size)
    t = torch.from_numpy(t_obs).to(gpu)
    w = torch.from_numpy(0.5 * w_true).to(gpu).requires_grad_(True)

    def loss_fn(ww):
        r = hmatvec(dev, ww) - t
        return (r * r).sum()

    loss = loss_fn(w)
    loss.backward()
    g = w.grad.detach().cpu().numpy()
    # against the oracle: 2 H^T (H w - t)
    w_np = 0.5 * w_true
    res = O.matvec_original(h.flat(), w_np) - t_obs
    g_ref = 2.0 * O.rmatvec_original(h.flat(), res)
    assert np.linalg.norm(g - g_ref) <= 1e-12 * np.linalg.norm(g_ref)
    # central finite differences on 20 random directions (quadratic loss:
    # exact up to rounding)
    with torch.no_grad():
        for k in range(20):
            v = torch.from_numpy(rng.standard_normal(h.size)).to(gpu)
            eps = 1e-3 * torch.linalg.vector_norm(w) / torch.linalg.vector_norm(v)
            fd = (loss_fn(w + eps * v) - loss_fn(w - eps * v)) / (2 * eps)
            an = (w.grad * v).sum()
            assert abs(fd.item() - an.item()) <= 1e-7 * max(abs(an.item()), 1e-30) + 1e-9 * loss.item()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_batched_gradient_matches_oracle(gpu, precision, tol):
    """Several observation sets at once: W is (n, 7), the forward pass is one
    batched H call and the backward one batched H^T call (MFMA tiles); the
    gradient matches 2 H^T (H W - T) from the oracle column by column."""
    import torch
    from bigwham_dev_amd.autograd import hmatvec
    coor, conn = rect_grid(24)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    dev = HipHmat(h, precisions=("fp64", "fp32"), adjoint=True)
    print(f"batched autograd {precision} n={h.size}", flush=True)
    rng = np.random.default_rng(3)
    W_np = rng.uniform(-1, 1, (h.size, 7))
    T_np = rng.uniform(-1, 1, (h.size, 7))
    W = torch.from_numpy(W_np).to(gpu).requires_grad_(True)
    T = torch.from_numpy(T_np).to(gpu)
    R = hmatvec(dev, W, precision=precision) - T
    (R * R).sum().backward()
    G = W.grad.cpu().numpy()
    f = h.flat()
    for j in range(7):
        res = O.matvec_original(f, W_np[:, j]) - T_np[:, j]
        g_ref = 2.0 * O.rmatvec_original(f, res)
        assert np.linalg.norm(G[:, j] - g_ref) <= tol * np.linalg.norm(g_ref), j
    # the adjoint batched call on its own
    Yt = dev.rmatvec_batched(T, precision=precision).cpu().numpy()
    for j in (0, 6):
        ref = O.rmatvec_original(f, T_np[:, j])
        assert np.linalg.norm(Yt[:, j] - ref) <= tol * np.linalg.norm(ref)
