"""GPU parity of the MI355X H-matvec against the CPU oracle (restatement of
Hmat::matvec / matvecOriginal, src/hmat/hmatrix/hmat.cc:472-715), through the
C-ABI of libbwhip.so.

Tolerances: FP64 1e-12 relative (summation order only, as the reference's own
OpenMP merge and FP64 atomics reorder sums); FP32 factors with FP32
accumulation 1e-5 relative."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HipHmat, HostHmat
from bigwham_dev_amd.meshes import rect_grid, segment_line

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def cases():
    out = {}
    h = HostHmat.gaussian(1500, 10.0)
    h.build(32, 3.0, 1e-6)
    out["gaussian"] = h
    coor, conn = segment_line(2000)
    h = HostHmat.from_mesh("2DS0-H", coor, conn, [1.0, 0.25])
    h.build(32, 3.0, 1e-6)
    out["2DS0"] = h
    coor, conn = rect_grid(24)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    out["3DR0"] = h
    h = HostHmat.from_mesh("3DR0-H-mode1", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    out["3DR0m1"] = h
    return out


@pytest.mark.parametrize("name", ["gaussian", "2DS0", "3DR0", "3DR0m1"])
@pytest.mark.parametrize("original", [False, True])
def test_matvec_fp64_fp32(gpu, cases, name, original):
    import torch
    h = cases[name]
    print(f"case {name} original={original} n={h.size}", flush=True)
    dev = HipHmat(h, precisions=("fp64", "fp32"))
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, h.size)
    y_ref = (O.matvec_original if original else O.matvec)(h.flat(), x)
    xd = torch.from_numpy(x).to(gpu)
    y = dev.matvec(xd, original=original).cpu().numpy()
    assert rel(y, y_ref) <= TOL64
    y2 = dev.matvec(xd, original=original).cpu().numpy()
    assert np.array_equal(y, y2), "matvec must be deterministic"
    yh = dev.matvec(x, original=original)  # host-pointer path
    assert np.array_equal(yh, y)
    y32 = dev.matvec(xd, original=original, precision="fp32").cpu().numpy()
    assert rel(y32, y_ref) <= TOL32


@pytest.mark.parametrize("name", ["gaussian", "2DS0", "3DR0"])
def test_rmatvec(gpu, cases, name):
    import torch
    h = cases[name]
    dev = HipHmat(h, precisions=("fp64", "fp32"), adjoint=True)
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, h.size)
    for original in (False, True):
        y_ref = (O.rmatvec_original if original else O.rmatvec)(h.flat(), x)
        y = dev.rmatvec(torch.from_numpy(x).to(gpu), original=original).cpu().numpy()
        assert rel(y, y_ref) <= TOL64
        y32 = dev.rmatvec(torch.from_numpy(x).to(gpu), original=original,
                          precision="fp32").cpu().numpy()
        assert rel(y32, y_ref) <= TOL32


def test_nb_entries(gpu, cases):
    """nbOfEntries (hmat.cc:299-312) as the device handle counts it, against
    the oracle's count of the same blocks and the host's own."""
    for name, h in cases.items():
        dev = HipHmat(h, precisions=("fp64", "fp32"))
        print(f"nb_entries {name}: {dev.nb_entries}", flush=True)
        assert dev.nb_entries == O.nb_entries(h.flat()) == h.nb_entries(), name
        # the streamed bytes hold every algorithmic entry once, plus padding
        assert dev.stream_bytes["fp64"] >= 8 * dev.nb_entries
        assert dev.stream_bytes["fp32"] >= 4 * dev.nb_entries


def test_diagonal(gpu, cases):
    for name, h in cases.items():
        dev = HipHmat(h)
        assert np.array_equal(dev.diagonal(original=False), O.diagonal(h.flat()))
        assert np.array_equal(dev.diagonal(original=True), O.diagonal(h.flat(), original=True))
    g = cases["gaussian"]
    assert HipHmat(g).diagonal().sum() == g.size


@pytest.mark.parametrize("nrhs", [1, 8, 32, 37])
@pytest.mark.parametrize("original", [False, True])
def test_batched_fp64_mfma(gpu, cases, nrhs, original):
    """Many right-hand sides in FP64 on MFMA (v_mfma_f64_16x16x4f64) against
    the oracle column by column, at the FP64 bar (1e-12)."""
    import torch
    for name in ("3DR0", "2DS0", "3DR0m1", "gaussian"):
        h = cases[name]
        print(f"fp64 batched case {name} nrhs={nrhs} original={original}", flush=True)
        dev = HipHmat(h, precisions=("fp64", "fp32"))
        rng = np.random.default_rng(100 + nrhs)
        X = rng.uniform(-1, 1, (h.size, nrhs))
        Y = dev.matvec_batched(torch.from_numpy(X).to(gpu), original=original).cpu().numpy()
        fn = O.matvec_original if original else O.matvec
        for k in range(nrhs):
            assert rel(Y[:, k], fn(h.flat(), X[:, k])) <= TOL64, (name, k)
        # two calls in a row give the same bits
        Y2 = dev.matvec_batched(torch.from_numpy(X).to(gpu), original=original).cpu().numpy()
        assert np.array_equal(Y, Y2)


@pytest.mark.parametrize("nrhs", [5, 32, 40])
@pytest.mark.parametrize("original", [False, True])
def test_batched_fp32_mfma(gpu, cases, nrhs, original):
    """Many right-hand sides on FP32 MFMA (v_mfma_f32_32x32x2f32) against the
    oracle, column by column; nrhs below, at and above the 32-wide tile."""
    import torch
    for name in ("3DR0", "2DS0", "gaussian"):
        h = cases[name]
        print(f"case {name} nrhs={nrhs} original={original}", flush=True)
        dev = HipHmat(h, precisions=("fp64", "fp32"))
        rng = np.random.default_rng(nrhs)
        X = rng.uniform(-1, 1, (h.size, nrhs))
        Y = dev.matvec_batched(torch.from_numpy(X).to(gpu), original=original,
                               precision="fp32").cpu().numpy()
        f = h.flat()
        for k in range(nrhs):
            ref = (O.matvec_original if original else O.matvec)(f, X[:, k])
            assert rel(Y[:, k], ref) <= TOL32, (name, k)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("nrhs", [6, 32, 34])
def test_batched_gather_paths_agree(gpu, cases, precision, nrhs):
    """A row-major X (gathered four columns per thread with 16-byte loads)
    and the same X stored column-major (gathered one element per thread) give
    the same bits; nrhs 6 and 34 leave partial four-column groups."""
    import torch
    h = cases["3DR0"]
    print(f"gather paths {precision} nrhs={nrhs}", flush=True)
    dev = HipHmat(h, precisions=("fp64", "fp32"))
    X = torch.from_numpy(np.random.default_rng(300 + nrhs).uniform(-1, 1, (h.size, nrhs))).to(gpu)
    Xc = X.t().contiguous().t()  # column-major copy of the same values
    assert X.stride() == (nrhs, 1) and Xc.stride() == (1, h.size)
    Yr = dev.matvec_batched(X, precision=precision)
    Yc = dev.matvec_batched(Xc, precision=precision)
    assert torch.equal(Yr, Yc)
    f = h.flat()
    tol = TOL32 if precision == "fp32" else TOL64
    for k in (0, nrhs - 1):
        assert rel(Yr[:, k].cpu().numpy(), O.matvec_original(f, X[:, k].cpu().numpy())) <= tol


@pytest.mark.parametrize("unroll", [4, 8, 16, (4, 16), (16, 8)])
def test_unroll_depths(gpu, cases, unroll):
    """Every unroll depth the autotuner may pick, the same in both passes or
    one per pass, keeps parity."""
    import torch
    h = cases["3DR0"]
    print(f"unroll {unroll} n={h.size}", flush=True)
    dev = HipHmat(h, precisions=("fp64", "fp32"))
    u1, u2 = unroll if isinstance(unroll, tuple) else (unroll, unroll)
    for p in ("fp64", "fp32"):
        dev.set_unroll_pass(p, 1, u1)
        dev.set_unroll_pass(p, 2, u2)
    x = np.random.default_rng(11).uniform(-1, 1, h.size)
    y_ref = O.matvec_original(h.flat(), x)
    xd = torch.from_numpy(x).to(gpu)
    assert rel(dev.matvec(xd).cpu().numpy(), y_ref) <= TOL64
    assert rel(dev.matvec(xd, precision="fp32").cpu().numpy(), y_ref) <= TOL32


def test_autotune_unroll(gpu, cases, tmp_path, monkeypatch):
    from bigwham_dev_amd import autotune
    monkeypatch.setenv("BIGWHAM_TUNE_CACHE", str(tmp_path / "tune.json"))
    dev = HipHmat(cases["3DR0"], precisions=("fp64", "fp32"))
    r = autotune.tune_unroll(dev, reps=2)
    assert set(r.unroll) == {"fp64", "fp32"}
    assert all(len(u) == 2 and set(u) <= set(autotune.UNROLLS) for u in r.unroll.values())
    r2 = autotune.tune_unroll(dev, reps=2)  # from the cache
    assert r2.unroll == r.unroll and not r2.ms


@pytest.mark.parametrize("name", ["3DR0", "gaussian"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batched_mfma_schedules(gpu, cases, name, precision):
    """Every schedule the autotuner may pick keeps parity (FP32 1e-5, FP64
    1e-12); the schedules of one MFMA shape issue the same MFMAs in the same
    order per accumulator, so their results agree bit for bit (FP32: 0-11 on
    32 x 32 tiles, 12-13 on 16 x 16 tiles, whose MFMAs sum four products
    instead of two; FP64: all on 16 x 16 tiles).  3DR0 has 64-row slabs (two
    32-row tiles), the Gaussian matrix (dof 1, leaf 32) slabs of at most 32
    rows; the small work items leave fewer steps than the deep schedules'
    ring holds."""
    import torch
    h = cases[name]
    dev = HipHmat(h, precisions=(precision,))
    X = torch.from_numpy(np.random.default_rng(13).uniform(-1, 1, (h.size, 32))).to(gpu)
    tol = TOL32 if precision == "fp32" else TOL64
    first = {}
    for schedule in range(16):
        print(f"case {name} {precision} mfma schedule {schedule}", flush=True)
        dev.set_mfma_schedule(schedule, precision)
        Y = dev.matvec_batched(X, precision=precision).cpu().numpy()
        for j in (0, 17, 31):
            ref = O.matvec_original(h.flat(), X[:, j].cpu().numpy())
            assert rel(Y[:, j], ref) <= tol, (schedule, j)
        shape = "16x16" if precision == "fp32" and schedule in (12, 13) else "main"
        if shape not in first:
            first[shape] = Y
        else:
            assert np.array_equal(Y, first[shape]), schedule


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batched_strides(gpu, cases, precision):
    """Row-major, column-major and sliced operands give the same bits as
    the column-major leading-dimension entry point, and bad strides fail."""
    import ctypes as C
    import torch
    from bigwham_dev_amd import _native as N
    h = cases["3DR0"]
    n = h.size
    print(f"strides {precision} n={n}", flush=True)
    dev = HipHmat(h, precisions=("fp64", "fp32"), adjoint=True)
    prec = N.BW_PREC_FP64 if precision == "fp64" else N.BW_PREC_FP32
    Xh = np.random.default_rng(21).uniform(-1, 1, (n, 20))
    Xr = torch.from_numpy(Xh).to(gpu)                     # row-major (20, 1)
    Xc = Xr.t().contiguous().t()                          # column-major (1, n)
    assert Xc.stride() == (1, n)
    # the leading-dimension entry point, column-major in and out
    Yld = torch.empty((20, n), dtype=torch.float64, device=gpu)
    stream = torch.cuda.current_stream(gpu).cuda_stream
    N.check_hip(N.hip().bw_hip_matvec_batched(
        dev._h, 20, C.c_void_p(Xc.data_ptr()), n, C.c_void_p(Yld.data_ptr()), n,
        N.BW_ORDER_ORIGINAL, prec, C.c_void_p(stream)))
    ref = Yld.t().cpu().numpy()
    Yr = dev.matvec_batched(Xr, precision=precision)
    assert Yr.stride() == (20, 1)
    assert np.array_equal(Yr.cpu().numpy(), ref)
    assert np.array_equal(dev.matvec_batched(Xc, precision=precision).cpu().numpy(), ref)
    # a column slice of a wider row-major matrix, into a column-major output
    Xw = torch.from_numpy(np.random.default_rng(22).uniform(-1, 1, (n, 30))).to(gpu)
    Xw[:, 5:25] = Xr
    out = torch.empty((20, n), dtype=torch.float64, device=gpu).t()
    Ys = dev.matvec_batched(Xw[:, 5:25], precision=precision, out=out)
    assert Ys.data_ptr() == out.data_ptr()
    assert np.array_equal(Ys.cpu().numpy(), ref)
    # the adjoint through the same entry point
    Yt = dev.rmatvec_batched(Xr, precision=precision).cpu().numpy()
    tol = TOL64 if precision == "fp64" else TOL32
    for j in (0, 19):
        assert rel(Yt[:, j], O.rmatvec_original(h.flat(), Xh[:, j])) <= tol
    # overlapping output strides and a zero stride are refused
    Ybad = torch.empty((n, 20), dtype=torch.float64, device=gpu)
    for ys in ((1, 1), (2, 1), (0, 1)):
        rc = N.hip().bw_hip_matvec_batched_strided(
            dev._h, 0, 20, C.c_void_p(Xr.data_ptr()), 20, 1, C.c_void_p(Ybad.data_ptr()),
            ys[0], ys[1], N.BW_ORDER_ORIGINAL, prec, C.c_void_p(stream))
        assert rc == N.BW_ESHAPE, ys
