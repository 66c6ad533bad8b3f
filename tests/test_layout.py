"""CPU checks of the device slab layout: bounds of every work item and a
numpy replay of the kernels against the oracle (no GPU needed)."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HostHmat
from layout_emu import LayoutDump
from bigwham_dev_amd.meshes import rect_grid, segment_line


def _cases():
    h = HostHmat.gaussian(700, 10.0)
    h.build(32, 3.0, 1e-6)
    yield "gaussian", h
    coor, conn = segment_line(900)
    h = HostHmat.from_mesh("2DS0-H", coor, conn, [1.0, 0.25])
    h.build(32, 3.0, 1e-6)
    yield "2DS0", h
    coor, conn = rect_grid(16)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    yield "3DR0", h


CASES = dict(_cases())


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("wi_bytes", [0, 2048])
def test_layout_replay(name, precision, adjoint, wi_bytes):
    h = CASES[name]
    f = h.flat()
    lay = LayoutDump(f, precision, adjoint, wi_bytes)
    lay.check_bounds()
    lay.check_xcd_cover()
    x = np.random.default_rng(3).uniform(-1, 1, h.size)
    y = lay.replay(x)
    y_ref = O.rmatvec(f, x) if adjoint else O.matvec(f, x)
    tol = 1e-13 if precision == "fp64" else 1e-6
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < tol
    perm = h.permutation()
    y = lay.replay(x, perm, perm)
    y_ref = O.rmatvec_original(f, x) if adjoint else O.matvec_original(f, x)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < tol


def test_layout_padding_and_split():
    """Work items of one slab are split at wi_bytes and reduced in order; the
    stored entries exceed the algorithmic ones only by column padding."""
    h = CASES["3DR0"]
    f = h.flat()
    lay = LayoutDump(f, "fp64", False, 2048)
    assert len(lay.passes[1]["reds"]) > 0
    stored = sum(P["data"].size for P in lay.passes)
    E = O.nb_entries(f)
    assert E <= stored <= 1.2 * E


def _meshes(ns, kernel="3DR0-H"):
    out = []
    for n in ns:
        coor, conn = rect_grid(n)
        h = HostHmat.from_mesh(kernel, coor, conn, [1.0, 0.25])
        h.build(64, 3.0, 1e-6)
        out.append(h)
    return out


MULTI = _meshes([10, 16, 7])


def _compose(hosts):
    import ctypes as C
    from bigwham_dev_amd import _native as N
    lib = N.hip()
    flats = [h.flat() for h in hosts]
    arr = (C.POINTER(N.FlatHmat) * len(flats))(*[C.pointer(f) for f in flats])
    out = C.c_void_p()
    rc = lib.bw_flat_compose(arr, len(flats), C.byref(out))
    return rc, out, flats


def test_composite_is_block_diagonal():
    """The composite of several meshes applies each mesh's own H to its slice:
    the oracle on the composite equals the per-mesh oracle results, bit for
    bit (same block order within each mesh), in both orderings."""
    import ctypes as C
    from bigwham_dev_amd import _native as N
    rc, c, flats = _compose(MULTI)
    assert rc == 0
    view = N.hip().bw_flat_composite_view(c).contents
    off = np.zeros(len(MULTI) + 1, np.int64)
    N.hip().bw_flat_composite_offsets(c, off.ctypes.data_as(C.POINTER(C.c_int64)), None)
    assert list(np.diff(off)) == [h.size for h in MULTI]
    x = np.random.default_rng(5).uniform(-1, 1, int(off[-1]))
    for fn in (O.matvec, O.matvec_original, O.rmatvec):
        y = fn(view, x)
        for m, f in enumerate(flats):
            sl = slice(off[m], off[m + 1])
            r = fn(f, x[sl])  # the oracle merges per-thread partial sums: order may differ
            assert np.abs(y[sl] - r).max() <= 1e-14 * np.abs(r).max()
    N.hip().bw_flat_composite_free(c)


def test_composite_rejects_mixed_dof():
    from bigwham_dev_amd import _native as N
    coor, conn = segment_line(200)
    h2 = HostHmat.from_mesh("2DS0-H", coor, conn, [1.0, 0.25])
    h2.build(32, 3.0, 1e-6)
    rc, c, _ = _compose([MULTI[0], h2])
    assert rc == N.BW_EINVAL and not c.value
    assert b"dof" in N.hip().bw_flat_compose_last_error()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("adjoint", [False, True])
def test_composite_layout_replay(precision, adjoint):
    flats = [h.flat() for h in MULTI]
    lay = LayoutDump(flats, precision, adjoint, 4096)
    lay.check_bounds()
    sizes = [h.size for h in MULTI]
    off = np.concatenate([[0], np.cumsum(sizes)])
    x = np.random.default_rng(9).uniform(-1, 1, off[-1])
    perm = np.concatenate([h.permutation() + off[m] // 3 for m, h in enumerate(MULTI)])
    fn = O.rmatvec_original if adjoint else O.matvec_original
    y_ref = np.concatenate([fn(f, x[off[m]:off[m + 1]]) for m, f in enumerate(flats)])
    y = lay.replay(x, perm, perm)
    tol = 1e-13 if precision == "fp64" else 1e-6
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < tol


@pytest.mark.parametrize("case", ["single", "fr_only", "fixed_rank", "homogeneous"])
def test_layout_edge_cases(case):
    """Layouts of degenerate patterns replay exactly like the oracle."""
    if case == "single":
        coor, conn = segment_line(1)
        h = HostHmat.from_mesh("2DS0-H", coor, conn, [1.0, 0.25])
        h.build(32, 3.0, 1e-6)
    elif case == "fr_only":
        coor, conn = rect_grid(5)
        h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
        h.build(64, 3.0, 1e-6)
    else:
        coor, conn = rect_grid(22)
        h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25],
                               fixed_rank=3 if case == "fixed_rank" else -1,
                               homogeneous_size_pattern=case == "homogeneous")
        h.build(32, 3.0, 1e-6)
    f = h.flat()
    for precision in ("fp64", "fp32"):
        lay = LayoutDump(f, precision, False, 1024)
        lay.check_bounds()
        x = np.random.default_rng(2).uniform(-1, 1, h.size)
        tol = 1e-13 if precision == "fp64" else 1e-6
        perm = h.permutation()
        y = lay.replay(x, perm, perm)
        y_ref = O.matvec_original(f, x)
        assert np.linalg.norm(y - y_ref) <= tol * np.linalg.norm(y_ref)


def test_xcd_ranges_balance_bytes():
    """Each XCD gets a contiguous range of work items of about 1/8 of the bytes."""
    coor, conn = rect_grid(30)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    lay = LayoutDump(h.flat(), "fp64", False, 1024)
    for p, P in enumerate(lay.passes):
        xs, _ = lay.xcd[p]
        w = P["wis"]
        assert len(w) > 64
        b = w[:, 2] * w[:, 3] + 1
        per = np.array([b[xs[x]:xs[x + 1]].sum() for x in range(8)])
        assert per.max() - per.min() <= 2 * b.max()


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("shape", [(150, 90, 3), (70, 260, 2), (200, 200, 1)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("adjoint", [False, True])
def test_rectangular_layout_replay(shape, precision, adjoint):
    """Receiver and source point sets of different sizes (random blocks,
    ranks 0-70): the oracle equals the numpy assembly of the same blocks, and
    the layout replay equals both, in both numberings."""
    from synthetic_hmat import SyntheticHmat
    S = SyntheticHmat(*shape, seed=sum(shape))
    f = S.flat()
    lay = LayoutDump(f, precision, adjoint, 4096)
    lay.check_bounds()
    lay.check_xcd_cover()
    x = np.random.default_rng(5).uniform(-1, 1, lay.n_in)
    tol = 1e-13 if precision == "fp64" else 1e-6
    for original in (False, True):
        H = S.dense(original)
        ref = H.T @ x if adjoint else H @ x
        if original:
            o = (O.rmatvec_original if adjoint else O.matvec_original)(f, x)
            pin, pout = (S.perm0, S.perm1) if adjoint else (S.perm1, S.perm0)
            y = lay.replay(x, pin, pout)
        else:
            o = (O.rmatvec if adjoint else O.matvec)(f, x)
            y = lay.replay(x)
        assert _rel(o, ref) < 1e-13
        assert _rel(y, ref) < tol


@pytest.mark.parametrize("shape", [(150, 90, 3), (70, 260, 2)])
def test_rectangular_oracle_diagonal(shape):
    """The oracle's diagonal of an operator between point sets of different
    sizes: dof * max(rows, cols) entries, each the dense operator's diagonal
    entry where a full-rank block crosses the diagonal and 0 elsewhere, and
    the original numbering is the permuted one moved by perm1 (more row than
    column points used to read perm1 past its end)."""
    from synthetic_hmat import SyntheticHmat
    S = SyntheticHmat(*shape, seed=sum(shape))
    f = S.flat()
    n = S.dof * max(S.n_row_pts, S.n_col_pts)
    dp = O.diagonal(f, original=False)
    do = O.diagonal(f, original=True)
    assert dp.shape == do.shape == (n,)
    H = S.dense(False)
    k = min(H.shape)
    on = dp[:k] != 0
    assert on.any()
    assert np.array_equal(dp[:k][on], np.diag(H)[on])
    assert not np.any(dp[k:])
    # original numbering: entry dof * perm1[i] + j holds permuted entry dof * i + j
    idx = (S.dof * np.asarray(S.perm1)[:, None] + np.arange(S.dof)[None, :]).ravel()
    expect = np.zeros(n)
    expect[idx] = dp[:idx.size]
    assert np.array_equal(do, expect)
