"""Pins of the CPU oracle (restatement of Hmat::matvec / matvecOriginal /
diagonal / nbOfEntries, src/hmat/hmatrix/hmat.cc:237-715) by brute force and
invariants, so that the oracle itself is checked before it checks the GPU."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HostHmat
from bigwham_dev_amd.meshes import penny_delaunay, rect_grid, segment_line


def _build(name):
    if name == "gaussian":
        h = HostHmat.gaussian(400, 10.0)
        h.build(32, 3.0, 1e-6)
    elif name == "2DS0":
        h = HostHmat.from_mesh("2DS0-H", *segment_line(500), [1.0, 0.25])
        h.build(32, 3.0, 1e-6)
    elif name == "3DR0":
        h = HostHmat.from_mesh("3DR0-H", *rect_grid(14), [1.0, 0.25])
        h.build(64, 3.0, 1e-6)
    elif name == "3DT0":
        h = HostHmat.from_mesh("3DT0-H", *penny_delaunay(300, seed=1), [1.0, 0.25])
        h.build(64, 3.0, 1e-6)
    return h


CASES = {k: _build(k) for k in ["gaussian", "2DS0", "3DR0", "3DT0"]}


def _perm_idx(perm, d):
    return (d * perm[:, None] + np.arange(d)[None, :]).ravel()


@pytest.mark.parametrize("name", list(CASES))
def test_matvec_equals_blockwise_dense(name):
    h = CASES[name]
    f = h.flat()
    Hd = O.dense(f)  # every FR block and every A*B placed at its pattern position
    # brute-force assembly in numpy from the per-block accessors
    d = h.dof
    Hb = np.zeros_like(Hd)
    fp, lp = h.fr_pattern(), h.lr_pattern()
    for k in range(f.n_fr):
        Hb[d * fp[1, k]:d * fp[3, k], d * fp[2, k]:d * fp[4, k]] = h.fr_block(k)
    for k in range(f.n_lr):
        A, B = h.lr_block(k)
        Hb[d * lp[1, k]:d * lp[3, k], d * lp[2, k]:d * lp[4, k]] = A @ B
    # FR blocks are copied exactly; A @ B differs from the oracle's loop only
    # by summation order
    assert np.max(np.abs(Hd - Hb)) <= 1e-14 * np.max(np.abs(Hb))
    x = np.random.default_rng(0).uniform(-1, 1, h.size)
    y = O.matvec(f, x)
    assert np.linalg.norm(y - Hd @ x) <= 1e-13 * np.linalg.norm(y)
    for nt in (1, 3, 8):
        assert np.linalg.norm(O.matvec(f, x, nt) - y) <= 1e-14 * np.linalg.norm(y)


@pytest.mark.parametrize("name", list(CASES))
def test_original_ordering_invariant(name):
    """matvecOriginal(x) = P0^T matvec(P1 x) (hmat.cc:586-626)."""
    h = CASES[name]
    f = h.flat()
    d = h.dof
    idx = _perm_idx(h.permutation(), d)
    x = np.random.default_rng(1).uniform(-1, 1, h.size)
    y = O.matvec(f, x[idx])
    yo = np.empty_like(y)
    yo[idx] = y
    assert np.array_equal(O.matvec_original(f, x, 1), O.matvec_original(f, x, 1))
    assert np.linalg.norm(O.matvec_original(f, x) - yo) <= 1e-14 * np.linalg.norm(yo)


@pytest.mark.parametrize("name", list(CASES))
def test_linearity_and_adjoint(name):
    h = CASES[name]
    f = h.flat()
    rng = np.random.default_rng(2)
    x, z = rng.uniform(-1, 1, (2, h.size))
    a, b = 0.7, -1.3
    lhs = O.matvec(f, a * x + b * z)
    rhs = a * O.matvec(f, x) + b * O.matvec(f, z)
    assert np.linalg.norm(lhs - rhs) <= 1e-13 * np.linalg.norm(rhs)
    Hd = O.dense(f)
    assert np.linalg.norm(O.rmatvec(f, z) - Hd.T @ z) <= 1e-13 * np.linalg.norm(Hd.T @ z)
    assert abs(O.matvec(f, x) @ z - x @ O.rmatvec(f, z)) <= 1e-12 * abs(x @ O.rmatvec(f, z))
    yo = O.rmatvec_original(f, z)
    Ho = h.dense(original=False)
    del Ho
    # original-order adjoint: <A_orig x, z> = <x, A_orig^T z>
    assert abs(O.matvec_original(f, x) @ z - x @ yo) <= 1e-12 * abs(x @ yo)


@pytest.mark.parametrize("name", list(CASES))
def test_diagonal_and_entries(name):
    h = CASES[name]
    f = h.flat()
    Hd = O.dense(f)
    assert np.array_equal(O.diagonal(f), np.diag(Hd))
    d = h.dof
    idx = _perm_idx(h.permutation(), d)
    dorig = np.empty(h.size)
    dorig[idx] = np.diag(Hd)
    assert np.array_equal(O.diagonal(f, original=True), dorig)
    assert O.nb_entries(f) == h.nb_entries()


def test_exchange_sort_literal():
    """oracle_exchange_sort is the literal double loop: sorted output, and
    on distinct keys the unique ascending order."""
    rng = np.random.default_rng(5)
    pts = rng.integers(0, 7, (300, 2)).astype(np.float64)
    perm = np.arange(300, dtype=np.int64)
    O.exchange_sort(pts, perm, 0, 300, 1)
    assert np.all(np.diff(pts[:, 1]) >= 0)
    pts2 = rng.uniform(size=(200, 3))
    p2 = np.arange(200, dtype=np.int64)
    orig = pts2.copy()
    O.exchange_sort(pts2, p2, 0, 200, 2)
    assert np.array_equal(p2, np.argsort(orig[:, 2]))
    assert np.array_equal(pts2, orig[p2])
