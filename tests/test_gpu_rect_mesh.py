"""Rectangular operators between two meshes on the device
(BEMatrixRectangular, the Python face of bigwham4py.BEMatrixRectangular,
bigwham4py.py:491-708): every kernel, a 2 000-6 000-element source mesh and a
different receiver mesh.  H x, H^T x and 32 right-hand sides in FP64 and
FP32 against the oracle's restatement of Hmat::matvecOriginal
(src/hmat/hmatrix/hmat.cc:572-635) on the same blocks, and the
LinearOperator semantics of the reference class."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd.bigwham4py import BEMatrixRectangular
from rect_meshes import KERNELS, edge_pairs, pair

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module", params=KERNELS)
def M(request, gpu):
    cs, ks, cr, kr, leaf = pair(request.param)
    print(f"rectangular {request.param}: {len(ks)} sources, {len(kr)} receivers", flush=True)
    m = BEMatrixRectangular(request.param, cs, ks, cr, kr, np.array([1.0, 0.25]),
                            max_leaf_size=leaf, eta=3.0, eps_aca=1e-6, verbose=False, device=0)
    yield m
    m.dev_.close()


def test_shape_and_helpers(M):
    h = M.H_
    nr, nc = h.shape_points()
    d = h.dof
    assert M.shape == (d * nr, d * nc) == (M.dev_.n_rows, M.dev_.n_cols)
    assert nr != nc
    assert sorted(M.getPermutation()) == list(range(nc))
    assert sorted(M.getPermutationReceivers()) == list(range(nr))
    assert M.getMeshCollocationPoints().shape == (nr, M.getSpatialDimension())
    assert 0 < M.getCompression() < 1
    assert M.dev_.nb_entries == O.nb_entries(h.flat())
    assert M._getPattern().shape[1] == 6
    assert M.getMatvecTime(3) > 0


def test_matvec_and_adjoint(M):
    f = M.H_.flat()
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, M.shape[1])
    u = rng.uniform(-1, 1, M.shape[0])
    ref = O.matvec_original(f, x)
    y = M @ x
    assert y.shape == (M.shape[0],)
    assert rel(y, ref) <= 1e-12
    assert np.array_equal(M @ x, y)  # repeatable bits
    reft = O.rmatvec_original(f, u)
    assert rel(M.T @ u, reft) <= 1e-12
    M.precision = "fp32"
    try:
        assert rel(M @ x, ref) <= 1e-5
        assert rel(M.T @ u, reft) <= 1e-5
    finally:
        M.precision = "fp64"


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batched_32(M, gpu, precision):
    import torch
    f = M.H_.flat()
    X = np.random.default_rng(8).uniform(-1, 1, (M.shape[1], 32))
    Y = M.dev_.matvec_batched(torch.from_numpy(X).to(gpu), original=True,
                              precision=precision).cpu().numpy()
    tol = 1e-12 if precision == "fp64" else 1e-5
    for j in range(0, 32, 5):
        assert rel(Y[:, j], O.matvec_original(f, X[:, j])) <= tol


def test_full_blocks(M):
    fb = M._getFullBlocks()
    assert fb.shape == M.shape
    assert fb.nnz == sum(
        (M.H_.dof * (p[3] - p[1])) * (M.H_.dof * (p[4] - p[2])) for p in M.H_.fr_pattern().T)
    # a few entries against the exact operator through the kernel helper
    d = M.H_.dof
    coo = fb.tocoo()
    if coo.nnz == 0:  # two parallel segment lines: every block is low rank
        return
    for k in np.random.default_rng(9).choice(coo.nnz, 20, replace=False):
        i, j = coo.row[k], coo.col[k]
        blk = M.H_.influence(j // d, i // d)
        assert blk[i % d, j % d] == coo.data[k]


@pytest.mark.parametrize("case", list(edge_pairs()))
def test_degenerate_sides(gpu, case):
    """One source element against a grid of receivers, and the reverse:
    single vectors and 5 / 32 right-hand sides, H and H^T, both precisions."""
    import torch
    kernel, cs, ks, cr, kr, leaf = edge_pairs()[case]
    print(f"rectangular {case}: {len(ks)} sources, {len(kr)} receivers", flush=True)
    m = BEMatrixRectangular(kernel, cs, ks, cr, kr, np.array([1.0, 0.25]), max_leaf_size=leaf,
                            eps_aca=1e-6, verbose=False, device=0)
    try:
        f = m.H_.flat()
        rng = np.random.default_rng(10)
        x = rng.uniform(-1, 1, m.shape[1])
        u = rng.uniform(-1, 1, m.shape[0])
        assert rel(m @ x, O.matvec_original(f, x)) <= 1e-12
        assert rel(m.T @ u, O.rmatvec_original(f, u)) <= 1e-12
        for k in (5, 32):
            for precision, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
                X = rng.uniform(-1, 1, (m.shape[1], k))
                Y = m.dev_.matvec_batched(torch.from_numpy(X).to(gpu), original=True,
                                          precision=precision).cpu().numpy()
                U = rng.uniform(-1, 1, (m.shape[0], k))
                V = m.dev_.rmatvec_batched(torch.from_numpy(U).to(gpu), original=True,
                                           precision=precision).cpu().numpy()
                for j in (0, k - 1):
                    assert rel(Y[:, j], O.matvec_original(f, X[:, j])) <= tol
                    assert rel(V[:, j], O.rmatvec_original(f, U[:, j])) <= tol
    finally:
        m.dev_.close()
