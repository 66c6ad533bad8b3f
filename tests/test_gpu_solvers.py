"""Device GMRES around the H-matvec on the GPU: FP64 and mixed-precision
solves reach an FP64-verified residual (checked with the CPU oracle of
Hmat::matvecOriginal, src/hmat/hmatrix/hmat.cc:572-715) and agree with a
dense direct solve of the oracle's dense H."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HipHmat, HostHmat, solvers
from bigwham_dev_amd.meshes import rect_grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    coor, conn = rect_grid(20)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-8)
    return h


def test_fp64_and_mixed_solves(gpu, system):
    import torch
    h = system
    dev = HipHmat(h, precisions=("fp64", "fp32"))
    print(f"solvers n={h.size}", flush=True)
    b_np = np.random.default_rng(2).uniform(-1, 1, h.size)
    b = torch.from_numpy(b_np).to(gpu)
    solvers.warm_up(dev, b)
    x64, i64 = solvers.solve_fp64(dev, b, tol=1e-10, restart=20)
    xm, im = solvers.solve_mixed(dev, b, tol=1e-10, restart=20)
    for x, info in ((x64, i64), (xm, im)):
        xn = x.cpu().numpy()
        r = np.linalg.norm(b_np - O.matvec_original(h.flat(), xn)) / np.linalg.norm(b_np)
        assert info.converged and r <= 1e-10
        assert info.rel_residual == pytest.approx(r, rel=1e-3)
    assert im.matvecs_fp64 <= 4 and im.matvecs_fp32 >= im.iterations
    # dense direct solve in the original ordering
    f = h.flat()
    P = h.permutation()
    D = O.dense(f)  # permuted ordering
    idx = (3 * P[:, None] + np.arange(3)).ravel()
    Dorig = np.empty_like(D)
    Dorig[np.ix_(idx, idx)] = D
    x_ref = np.linalg.solve(Dorig, b_np)
    for x in (x64, xm):
        assert np.linalg.norm(x.cpu().numpy() - x_ref) / np.linalg.norm(x_ref) < 1e-8


def test_pipelined_gmres_matches_host_order(gpu, system):
    """The look-ahead (pipelined) GMRES on the GPU follows the same iterates
    as the synchronous one on the CPU with the same operator."""
    import torch
    h = system
    dev = HipHmat(h, precisions=("fp64",))
    b_np = np.random.default_rng(4).uniform(-1, 1, h.size)
    b = torch.from_numpy(b_np).to(gpu)
    x_gpu, its_gpu, _ = solvers.gmres(lambda v: dev.matvec(v), b, None, tol=1e-9, restart=15)
    f = h.flat()
    x_cpu, its_cpu, _ = solvers.gmres(
        lambda v: torch.from_numpy(O.matvec_original(f, v.numpy())), torch.from_numpy(b_np),
        None, tol=1e-9, restart=15)
    assert abs(its_gpu - its_cpu) <= 1
    xg = x_gpu.cpu().numpy()
    assert np.linalg.norm(xg - x_cpu.numpy()) / np.linalg.norm(x_cpu.numpy()) < 1e-8


@pytest.mark.parametrize("tol", [1e-7, 1e-6])
def test_recompressed_fp32_copy(gpu, tol):
    """The FP32 copy with recompressed low-rank blocks
    (bw_hip_set_fp32_rank_tolerance): its matvec, in both directions, stays
    within 1e-5 of the oracle on the original blocks while streaming fewer
    entries; the FP64 copy is unchanged (1e-12); mixed GMRES on it reaches an
    FP64-verified 1e-8 (residual from the oracle)."""
    import torch
    from bigwham_dev_amd.meshes import penny_delaunay
    h = HostHmat.from_mesh("3DT0-H", *penny_delaunay(4000, seed=1), [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    f = h.flat()
    print(f"recompressed fp32 copy n={h.size} tol={tol}", flush=True)
    dev = HipHmat(h, precisions=("fp64", "fp32"), adjoint=True, fp32_rank_tol=tol)
    e64, e32 = dev.entries["fp64"], dev.entries["fp32"]
    assert sum(e64[:3]) == O.nb_entries(f) == dev.nb_entries
    assert e32[0] == e64[0] and e32[1] + e32[2] < e64[1] + e64[2]
    assert dev.stream_bytes["fp32"] < 0.5 * dev.stream_bytes["fp64"]
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, h.size)
    xd = torch.from_numpy(x).to(gpu)
    ref = O.matvec_original(f, x)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(dev.matvec(xd).cpu().numpy(), ref) <= 1e-12
    assert rel(dev.matvec(xd, precision="fp32").cpu().numpy(), ref) <= 1e-5
    reft = O.rmatvec_original(f, x)
    assert rel(dev.rmatvec(xd, precision="fp32").cpu().numpy(), reft) <= 1e-5
    X = torch.from_numpy(rng.uniform(-1, 1, (h.size, 32))).to(gpu)
    Y = dev.matvec_batched(X, precision="fp32").cpu().numpy()
    for j in (0, 31):
        assert rel(Y[:, j], O.matvec_original(f, X[:, j].cpu().numpy())) <= 1e-5
    b = torch.zeros(h.size, dtype=torch.float64, device=gpu)
    b[2::3] = 1.0  # unit normal pressure
    solvers.warm_up(dev, b)
    xm, im = solvers.solve_mixed(dev, b, tol=1e-8)
    r = np.linalg.norm(b.cpu().numpy() - O.matvec_original(f, xm.cpu().numpy()))
    r /= np.linalg.norm(b.cpu().numpy())
    assert im.converged and r <= 1e-8, r
