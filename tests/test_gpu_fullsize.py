"""Parity at the benchmark sizes: the 3DR0-H 200x200 square crack
(120 000 DoF), the 3DR0-H-mode1 500x500 crack (250 000 DoF) and the 3DT0-H
penny crack of ~200 000 triangles (599 739 DoF), FP64 against the oracle
(restatement of Hmat::matvecOriginal, src/hmat/hmatrix/hmat.cc:572-715) at
1e-12, FP32 at 1e-5, plus size-independent properties: the adjoint identity,
linearity and bitwise determinism; the gradient of ||H w - t||^2 through
torch.autograd against the oracle's 2 H^T (H w - t) and central differences;
the penny's FP64 and mixed-precision GMRES to 1e-8 with the residual taken
by the oracle; and the 3DR0-H 224x224 crack (150 528 DoF) with 32
right-hand sides through the batched MFMA path."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HipHmat, HostHmat
from bigwham_dev_amd.meshes import penny_delaunay, rect_grid

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("kernel,n", [("3DR0-H", 200), ("3DR0-H-mode1", 500)])
def test_full_size(gpu, kernel, n):
    import torch
    print(f"full size {kernel} {n}x{n}", flush=True)
    coor, conn = rect_grid(n)
    h = HostHmat.from_mesh(kernel, coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    print(f"built N={h.size} E={h.nb_entries()}", flush=True)
    dev = HipHmat(h, precisions=("fp64", "fp32"), adjoint=True)
    f = h.flat()
    rng = np.random.default_rng(1234)
    x = rng.uniform(-1, 1, h.size)
    ref = O.matvec_original(f, x)
    xd = torch.from_numpy(x).to(gpu)
    y = dev.matvec(xd)
    assert rel(y.cpu().numpy(), ref) <= 1e-12
    assert torch.equal(y, dev.matvec(xd))
    assert rel(dev.matvec(xd, precision="fp32").cpu().numpy(), ref) <= 1e-5
    w = torch.from_numpy(rng.uniform(-1, 1, h.size)).to(gpu)
    lhs = float(torch.dot(y, w))
    rhs = float(torch.dot(xd, dev.rmatvec(w)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    y2 = dev.matvec(2.0 * xd - 3.0 * w)
    assert float(torch.linalg.vector_norm(y2 - (2.0 * y - 3.0 * dev.matvec(w))) /
                 torch.linalg.vector_norm(y2)) <= 1e-13
    # gradient of ||H w - t||^2 through the adjoint kernel: the oracle's
    # 2 H^T (H w - t) at 1e-12, and central differences in 20 random
    # directions (the loss is quadratic, so they are exact up to rounding)
    from bigwham_dev_amd.autograd import hmatvec
    t_np = ref + 0.01 * np.abs(ref).mean() * rng.standard_normal(h.size)
    w_np = 0.5 * x
    td = torch.from_numpy(t_np).to(gpu)
    wv = torch.from_numpy(w_np).to(gpu).requires_grad_(True)

    def loss_fn(ww):
        r = hmatvec(dev, ww) - td
        return (r * r).sum()

    loss = loss_fn(wv)
    loss.backward()
    g_ref = 2.0 * O.rmatvec_original(f, O.matvec_original(f, w_np) - t_np)
    assert rel(wv.grad.cpu().numpy(), g_ref) <= 1e-12
    with torch.no_grad():
        for _ in range(20):
            v = torch.from_numpy(rng.standard_normal(h.size)).to(gpu)
            eps = 1e-3 * torch.linalg.vector_norm(wv) / torch.linalg.vector_norm(v)
            fd = (loss_fn(wv + eps * v) - loss_fn(wv - eps * v)) / (2 * eps)
            an = (wv.grad * v).sum()
            assert abs(fd.item() - an.item()) <= 1e-7 * abs(an.item()) + 1e-9 * loss.item()


def test_full_size_penny(gpu):
    """The 600k-DoF 3DT0-H penny matrix (FP64 copy and the FP32 copy with its
    low-rank blocks recompressed to 1e-7, as benchmarked: 76 GB of HBM); then
    the benchmarked solves: GMRES to 1e-8 with the Jacobi preconditioner,
    FP64 and mixed precision, each checked with the oracle's FP64 matvec."""
    import torch
    from bigwham_dev_amd import solvers
    print("full size 3DT0-H penny, 105 000 points", flush=True)
    coor, conn = penny_delaunay(105000, seed=0)
    h = HostHmat.from_mesh("3DT0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    print(f"built N={h.size} E={h.nb_entries()}", flush=True)
    dev = HipHmat(h, precisions=("fp64", "fp32"), fp32_rank_tol=1e-7)
    print("uploaded (FP32 copy recompressed)", flush=True)
    rng = np.random.default_rng(1234)
    x = rng.uniform(-1, 1, h.size)
    ref = O.matvec_original(h.flat(), x)
    xd = torch.from_numpy(x).to(gpu)
    y = dev.matvec(xd)
    assert rel(y.cpu().numpy(), ref) <= 1e-12
    assert torch.equal(y, dev.matvec(xd))
    assert rel(dev.matvec(xd, precision="fp32").cpu().numpy(), ref) <= 1e-5
    w = torch.from_numpy(rng.uniform(-1, 1, h.size)).to(gpu)
    y2 = dev.matvec(2.0 * xd - 3.0 * w)
    assert float(torch.linalg.vector_norm(y2 - (2.0 * y - 3.0 * dev.matvec(w))) /
                 torch.linalg.vector_norm(y2)) <= 1e-13
    b = torch.zeros(h.size, dtype=torch.float64, device=gpu)
    b[h.dof - 1::h.dof] = 1.0  # unit normal pressure (normal component last)
    bn = b.cpu().numpy()
    x64, i64 = solvers.solve_fp64(dev, b, tol=1e-8)
    xm, im = solvers.solve_mixed(dev, b, tol=1e-8)
    print(f"gmres fp64 {i64.iterations} its, mixed {im.matvecs_fp32} fp32 + "
          f"{im.matvecs_fp64} fp64 matvecs", flush=True)
    for sol, info in ((x64, i64), (xm, im)):
        r = rel(O.matvec_original(h.flat(), sol.cpu().numpy()), bn)
        assert info.converged and r <= 1e-8, r
    assert float(torch.linalg.vector_norm(xm - x64) / torch.linalg.vector_norm(x64)) <= 1e-6


def test_full_size_batched(gpu):
    """32 right-hand sides on the 224x224 3DR0-H crack, FP32 and FP64 MFMA,
    against the oracle column by column (first, middle and last columns)."""
    import torch
    print("full size batched 3DR0-H 224x224 x 32", flush=True)
    coor, conn = rect_grid(224)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    dev = HipHmat(h, precisions=("fp64", "fp32"))
    X = np.random.default_rng(7).uniform(-1, 1, (h.size, 32))
    Xd = torch.from_numpy(X).to(gpu)
    Y32 = dev.matvec_batched(Xd, precision="fp32").cpu().numpy()
    Y64 = dev.matvec_batched(Xd, precision="fp64").cpu().numpy()
    for j in (0, 15, 31):
        ref = O.matvec_original(h.flat(), X[:, j])
        assert rel(Y64[:, j], ref) <= 1e-12, j
        assert rel(Y32[:, j], ref) <= 1e-5, j
