"""The BEMatrix mirror (bigwham_dev_amd.bigwham4py; reference
interfaces/python/bigwham4py/bigwham4py.py:49-485) on the GPU engine: the
LinearOperator contract, the preconditioner helpers, selection by indexing
and the device solvers, against the oracle."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd.bigwham4py import BEMatrix
from bigwham_dev_amd.meshes import rect_grid

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def M():
    coor, conn = rect_grid(20)
    return BEMatrix("3DR0-H", coor, conn, np.array([1.0, 0.25]), max_leaf_size=64, eta=3.0,
                    eps_aca=1e-6, verbose=False)


def test_linear_operator(gpu, M):
    n = M.shape[0]
    assert M.shape == (n, n) and M.dtype == np.float64
    f = M.H_.flat()
    x = np.random.default_rng(0).uniform(-1, 1, n)
    assert rel(M @ x, O.matvec_original(f, x)) <= 1e-12
    assert rel(M.rmatvec(x), O.rmatvec_original(f, x)) <= 1e-12
    assert 0 < M.getCompression() < 1
    assert np.array_equal(M.H_diag(), O.diagonal(f, original=True))
    assert M.getSpatialDimension() == 3
    pat = M._getPattern()
    assert pat.shape[1] == 6 and set(np.unique(pat[:, 4])) <= {0, 1}


def test_full_blocks_and_ilu(gpu, M):
    fb = M._getFullBlocks()
    assert np.array_equal(fb.diagonal(), M.H_diag())
    half = M._getFullBlocks(keep_ratio=0.5)
    assert half.nnz <= fb.nnz // 2 + M.shape[0]
    P = M.H_ILU_prec(fill_factor=2, drop_tol=1e-4)
    x = np.random.default_rng(1).uniform(-1, 1, M.shape[0])
    z = P @ (M @ x)
    assert np.all(np.isfinite(z)) and rel(z, x) < 0.9  # the near field dominates


def test_scipy_gmres_with_jacobi(gpu, M):
    from scipy.sparse.linalg import gmres
    b = np.ones(M.shape[0])
    x, info = gmres(M, b, M=M.H_jacobi_prec(), rtol=1e-10, restart=30, maxiter=500)
    assert info == 0
    assert rel(M @ x, b) < 1e-9


def test_device_solve(gpu, M):
    import torch
    b = torch.ones(M.shape[0], dtype=torch.float64, device=gpu)
    for mixed in (False, True):
        x, info = M.solve(b, tol=1e-10, mixed=mixed)
        assert info.converged and info.rel_residual <= 1e-10


def test_selection_by_indexing(gpu, M):
    d = 3
    rng = np.random.default_rng(3)
    rows = rng.choice(M.shape[0] // d, 50, replace=False)
    cols = rng.choice(M.shape[0] // d, 70, replace=False)
    S = M[np.ix_(rows, cols)]
    assert S.shape == (d * rows.size, d * cols.size)
    x = rng.uniform(-1, 1, S.shape[1])
    assert rel(S @ x, O.selection_matvec(M.H_.flat(), rows, cols, x)) <= 1e-12
    fb = S._getFullBlocks()
    assert fb.shape == S.shape and fb.nnz > 0
    with pytest.raises(NotImplementedError):
        S.H_diag()


def test_write_hmatrix_replays(gpu, M, tmp_path):
    """write_hmatrix saves the operator; the saved copy replays bit for bit."""
    import torch
    from bigwham_dev_amd import HipHmat, store
    assert M.write_hmatrix(str(tmp_path / "h")) == 0
    saved = store.load(tmp_path / "h")
    dev = HipHmat(saved, precisions=("fp64",))
    x = torch.from_numpy(np.random.default_rng(9).uniform(-1, 1, M.shape[0])).to(gpu)
    assert torch.equal(dev.matvec(x), M.dev_.matvec(x))
    assert M.isCudaAvailable() and M.get_omp_threads() >= 1
