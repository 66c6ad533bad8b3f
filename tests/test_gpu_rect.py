"""Rectangular operators on the device: receiver and source point sets of
different sizes, as the reference's Hmat between two meshes, with random
blocks of ranks 0 to 70 (synthetic_hmat.py).  H x, H^T x, FP32 factors and the
batched MFMA path in both numberings, against the oracle (restatement of
Hmat::matvec / matvecOriginal, src/hmat/hmatrix/hmat.cc:472-715) and the
numpy assembly of the same blocks."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HipHmat
from synthetic_hmat import SyntheticHmat

pytestmark = pytest.mark.gpu

SHAPES = [(150, 90, 3), (70, 260, 2)]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module", params=SHAPES, ids=lambda s: f"{s[0]}x{s[1]}d{s[2]}")
def rect(request, gpu):
    S = SyntheticHmat(*request.param, seed=sum(request.param))
    print(f"rectangular {request.param}", flush=True)
    dev = HipHmat(S, precisions=("fp64", "fp32"), adjoint=True, device=0)
    assert (dev.n_rows, dev.n_cols) == (S.dof * S.n_row_pts, S.dof * S.n_col_pts)
    yield S, dev
    dev.close()


@pytest.mark.parametrize("original", [False, True])
def test_rect_matvec_and_adjoint(rect, gpu, original):
    import torch
    S, dev = rect
    f = S.flat()
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, dev.n_cols)
    ref = (O.matvec_original if original else O.matvec)(f, x)
    assert rel(ref, S.dense(original) @ x) < 1e-13
    xd = torch.from_numpy(x).to(gpu)
    assert rel(dev.matvec(xd, original=original).cpu().numpy(), ref) <= 1e-12
    assert rel(dev.matvec(xd, original=original, precision="fp32").cpu().numpy(), ref) <= 1e-5
    u = rng.uniform(-1, 1, dev.n_rows)
    reft = (O.rmatvec_original if original else O.rmatvec)(f, u)
    assert rel(reft, S.dense(original).T @ u) < 1e-13
    ud = torch.from_numpy(u).to(gpu)
    assert rel(dev.rmatvec(ud, original=original).cpu().numpy(), reft) <= 1e-12
    assert rel(dev.rmatvec(ud, original=original, precision="fp32").cpu().numpy(), reft) <= 1e-5


@pytest.mark.parametrize("nrhs", [5, 32])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("adjoint", [False, True])
def test_rect_batched(rect, gpu, nrhs, precision, adjoint):
    import torch
    S, dev = rect
    H = S.dense(True)
    H = H.T if adjoint else H
    X = np.random.default_rng(2).uniform(-1, 1, (H.shape[1], nrhs))
    ref = H @ X
    Y = dev.matvec_batched(torch.from_numpy(X).to(gpu), precision=precision,
                           adjoint=adjoint).cpu().numpy()
    tol = 1e-12 if precision == "fp64" else 1e-5
    for j in range(nrhs):
        assert rel(Y[:, j], ref[:, j]) <= tol


@pytest.mark.parametrize("original", [False, True])
def test_rect_selection(rect, gpu, original):
    """H[rows, cols] of a rectangular operator (bw_hip_create_selection):
    rows index the receiver points, columns the source points."""
    import torch
    S, _ = rect
    d = S.dof
    rng = np.random.default_rng(3)
    rows = rng.choice(S.n_row_pts, S.n_row_pts // 3, replace=False)
    cols = rng.choice(S.n_col_pts, S.n_col_pts // 2, replace=False)
    dev = HipHmat.selection(S, rows, cols, original=original, precisions=("fp64", "fp32"),
                            adjoint=True, device=0)
    assert (dev.n_rows, dev.n_cols) == (d * rows.size, d * cols.size)
    ri = (d * rows[:, None] + np.arange(d)).ravel()
    ci = (d * cols[:, None] + np.arange(d)).ravel()
    Hs = S.dense(original)[np.ix_(ri, ci)]
    x = rng.uniform(-1, 1, d * cols.size)
    ref = O.selection_matvec(S.flat(), rows, cols, x, original=original)
    assert rel(ref, Hs @ x) < 1e-13
    xd = torch.from_numpy(x).to(gpu)
    assert rel(dev.matvec(xd, original=original).cpu().numpy(), ref) <= 1e-12
    assert rel(dev.matvec(xd, original=original, precision="fp32").cpu().numpy(), ref) <= 1e-5
    u = rng.uniform(-1, 1, d * rows.size)
    reft = O.selection_matvec(S.flat(), rows, cols, u, original=original, adjoint=True)
    assert rel(reft, Hs.T @ u) < 1e-13
    assert rel(dev.rmatvec(torch.from_numpy(u).to(gpu), original=original).cpu().numpy(),
               reft) <= 1e-12
    dev.close()


def test_rect_composite(gpu):
    """Rectangular operators of different shapes as one block-diagonal
    operator (bw_hip_create_multi): each output slice equals its own
    operator's product."""
    import torch
    mats = [SyntheticHmat(150, 90, 3, seed=21), SyntheticHmat(40, 130, 3, seed=22),
            SyntheticHmat(95, 95, 3, seed=23)]
    dev = HipHmat.many(mats, precisions=("fp64", "fp32"), adjoint=True, device=0)
    ro, co = dev.row_offsets, dev.col_offsets
    assert ro[-1] == dev.n_rows and co[-1] == dev.n_cols
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, dev.n_cols)
    u = rng.uniform(-1, 1, dev.n_rows)
    for original in (False, True):
        y = dev.matvec(torch.from_numpy(x).to(gpu), original=original).cpu().numpy()
        yt = dev.rmatvec(torch.from_numpy(u).to(gpu), original=original).cpu().numpy()
        for k, S in enumerate(mats):
            H = S.dense(original)
            assert rel(y[ro[k]:ro[k + 1]], H @ x[co[k]:co[k + 1]]) <= 1e-12
            assert rel(yt[co[k]:co[k + 1]], H.T @ u[ro[k]:ro[k + 1]]) <= 1e-12
    dev.close()


def test_rect_diagonal(rect, gpu):
    """bw_hip_diagonal writes dof * max(n_rows, n_cols) entries (the length
    Hmat::diagonal returns, hmat.cc:256-285): the FR entries on the diagonal,
    zero past the shorter side, in both numberings, as the oracle."""
    S, dev = rect
    f = S.flat()
    n = S.dof * max(S.n_row_pts, S.n_col_pts)
    for original in (False, True):
        d = dev.diagonal(original=original)
        assert d.shape == (n,)
        assert np.array_equal(d, O.diagonal(f, original=original))
    dp = dev.diagonal(original=False)
    H = S.dense(False)
    k = min(H.shape)
    # entries are either the operator's diagonal (an FR block crosses it) or 0
    on = dp[:k] != 0
    assert np.allclose(dp[:k][on], np.diag(H)[on], rtol=0, atol=1e-13 * np.abs(H).max())
    assert not np.any(dp[k:])
