"""Low-rank recompression of the FP32 copy (bw_flat_recompress, the host step
behind bw_hip_set_fp32_rank_tolerance), pinned against LAPACK's SVD of each
block's product A B (numpy): the new rank is the smallest k whose discarded
singular values have a root sum of squares <= tol ||A B||_F, and the new
factors reproduce A B to that tolerance.  Blocks whose rank does not drop keep
their factors.  The operator built from the recompressed blocks stays within
the tolerance of the original (oracle matvec on both)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HostHmat
from bigwham_dev_amd import _native as N
from bigwham_dev_amd.meshes import penny_delaunay, rect_grid


@pytest.fixture(scope="module", params=["3DR0", "3DT0"])
def host(request):
    if request.param == "3DR0":
        h = HostHmat.from_mesh("3DR0-H", *rect_grid(24), [1.0, 0.25])
    else:
        h = HostHmat.from_mesh("3DT0-H", *penny_delaunay(900, seed=3), [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    return h


def recompress(host, tol):
    out = C.c_void_p()
    rc = N.hip().bw_flat_recompress(C.byref(host.flat()), tol, C.byref(out))
    if rc != N.BW_OK:
        raise N.BigWhamError(rc, N.hip().bw_flat_recompress_last_error().decode())
    return out


def blocks(flat, k, d):
    p = np.ctypeslib.as_array(flat.lr_pattern, shape=(flat.n_lr, 6))[k]
    m, n, r = d * (p[3] - p[1]), d * (p[4] - p[2]), int(p[5])
    if r == 0:
        return np.zeros((m, 0)), np.zeros((0, n))
    A = np.ctypeslib.as_array(flat.lr_A[k], shape=(r, m)).T
    B = np.ctypeslib.as_array(flat.lr_B[k], shape=(n, r)).T
    return A, B


@pytest.mark.parametrize("tol", [1e-7, 1e-6])
def test_ranks_and_errors_match_svd(host, tol):
    lib = N.hip()
    r = recompress(host, tol)
    try:
        new = lib.bw_flat_recompressed_view(r).contents
        old = host.flat()
        d = old.dof
        assert new.n_lr == old.n_lr and new.n_fr == old.n_fr
        before, after, changed = C.c_int64(), C.c_int64(), C.c_int64()
        lib.bw_flat_recompressed_stats(r, C.byref(before), C.byref(after), C.byref(changed))
        assert after.value < before.value and changed.value > 0
        n_exact = n_checked = 0
        e_after = 0
        for k in range(old.n_lr):
            A, B = blocks(old, k, d)
            A2, B2 = blocks(new, k, d)
            e_after += A2.size + B2.size
            if A.shape[1] == 0:
                continue
            M = A @ B
            s = np.linalg.svd(M, compute_uv=False)
            nrm = np.sqrt(np.sum(s ** 2))
            tail = np.sqrt(np.cumsum((s ** 2)[::-1]))[::-1]
            k_svd = int(np.argmax(np.append(tail, 0.0) <= tol * nrm))
            k_new = A2.shape[1]
            if k_new == A.shape[1]:
                # unchanged: same factors (the recompression gained nothing)
                assert new.lr_A[k] == old.lr_A[k] or np.array_equal(A2, A)
                assert k_svd >= k_new - 1
                continue
            n_checked += 1
            n_exact += k_new == k_svd
            assert abs(k_new - k_svd) <= 1, (k, k_new, k_svd)
            err = np.linalg.norm(M - A2 @ B2) / nrm
            assert err <= tol * 1.01 + 1e-12, (k, err)
        assert e_after == after.value  # low-rank entries only
        assert n_checked > 0 and n_exact >= 0.95 * n_checked
    finally:
        lib.bw_flat_recompressed_free(r)


def test_operator_error(host):
    """The whole operator moves by at most the block tolerance."""
    r = recompress(host, 1e-6)
    try:
        new = N.hip().bw_flat_recompressed_view(r).contents
        x = np.random.default_rng(4).uniform(-1, 1, host.size)
        y = O.matvec_original(host.flat(), x)
        y2 = O.matvec_original(new, x)
        assert np.linalg.norm(y2 - y) <= 1e-6 * np.linalg.norm(y)
        assert O.nb_entries(new) < O.nb_entries(host.flat())
    finally:
        N.hip().bw_flat_recompressed_free(r)


def test_bad_tolerance(host):
    out = C.c_void_p()
    for tol in (0.0, -1.0, 1.0, float("nan")):
        assert N.hip().bw_flat_recompress(C.byref(host.flat()), tol, C.byref(out)) == N.BW_EINVAL
    assert N.hip().bw_hip_set_fp32_rank_tolerance(-0.5) == N.BW_EINVAL
    assert N.hip().bw_hip_set_fp32_rank_tolerance(0.0) == N.BW_OK
