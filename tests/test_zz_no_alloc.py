"""bw_hip_matvec_noalloc: the matvec whose scratch is one fixed device array
per GPU, so that it never allocates.  It has no counterpart in the
reference; it is checked against bw_hip_matvec (same kernels, same order,
hence the same bits) through the C-ABI, on two calls in a row of which the
second is smaller (the array then holds the larger call's values), in both
precisions and orderings, and against the oracle (restatement of
Hmat::matvecOriginal / Hmat::matvec, src/hmat/hmatrix/hmat.cc:472-715)."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HipHmat, HostHmat
from bigwham_dev_amd import _native as N
from bigwham_dev_amd.meshes import rect_grid, segment_line

pytestmark = pytest.mark.gpu

ORDERS = {"original": N.BW_ORDER_ORIGINAL, "permuted": N.BW_ORDER_PERMUTED}
PRECS = {"fp64": N.BW_PREC_FP64, "fp32": N.BW_PREC_FP32}
TOL = {"fp64": 1e-12, "fp32": 1e-5}


@pytest.fixture(scope="module")
def ops(gpu):
    big = HostHmat.from_mesh("3DR0-H", *rect_grid(40), [1.0, 0.25])  # 4 800 DoF
    big.build(64, 3.0, 1e-6)
    small = HostHmat.from_mesh("2DS0-H", *segment_line(300), [1.0, 0.25])  # 600 DoF
    small.build(32, 3.0, 1e-6)
    devs = [HipHmat(h, precisions=("fp64", "fp32"), device=0) for h in (big, small)]
    yield list(zip((big, small), devs))
    for d in devs:
        d.close()


def _noalloc(dev, x, y, order, prec, stream=0):
    return N.hip().bw_hip_matvec_noalloc(dev._h, C.c_void_p(x.data_ptr()),
                                         C.c_void_p(y.data_ptr()), order, prec,
                                         C.c_void_p(stream))


def _oracle(host, x, order):
    f = host.flat()
    return O.matvec_original(f, x) if order == N.BW_ORDER_ORIGINAL else O.matvec(f, x)


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("order", list(ORDERS))
def test_noalloc_matches_matvec(ops, gpu, prec, order):
    import torch
    rng = np.random.default_rng(11)
    for rep in range(2):  # big then small, twice: each call follows a larger or other one
        for host, dev in ops:
            n = host.size
            print(f"noalloc {prec} {order} rep {rep} n={n}", flush=True)
            x = rng.uniform(-1, 1, n)
            xd = torch.from_numpy(x).to(gpu)
            y_na = torch.full((n,), np.nan, dtype=torch.float64, device=gpu)
            stream = torch.cuda.current_stream(gpu).cuda_stream
            rc = _noalloc(dev, xd, y_na, ORDERS[order], PRECS[prec], stream)
            assert rc == 0, N.hip().bw_hip_last_error().decode()
            y_ref = dev.matvec(xd, original=(order == "original"), precision=prec)
            torch.cuda.synchronize()
            assert torch.equal(y_na, y_ref), (prec, order, n)
            ref = _oracle(host, x, ORDERS[order])
            got = y_na.cpu().numpy()
            assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= TOL[prec]


def test_noalloc_allocates_nothing(ops, gpu):
    """A call from a stream the handle has never seen leaves the free device
    memory unchanged; bw_hip_matvec on that stream allocates its scratch
    there (the measurement sees allocations)."""
    import torch
    host, dev = ops[0]
    n = host.size
    x = torch.from_numpy(np.random.default_rng(5).uniform(-1, 1, n)).to(gpu)
    y = torch.empty(n, dtype=torch.float64, device=gpu)
    y2 = torch.empty_like(y)
    s2 = torch.cuda.Stream(gpu)
    with torch.cuda.stream(s2):
        y2.zero_()  # the stream's first launch, before the measurement
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(0)[0]
    assert _noalloc(dev, x, y, N.BW_ORDER_ORIGINAL, N.BW_PREC_FP64, s2.cuda_stream) == 0
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info(0)[0] == free0
    dev.matvec_ptr(x.data_ptr(), y2.data_ptr(), True, "fp64", s2.cuda_stream)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info(0)[0] < free0
    assert torch.equal(y, y2)
    N.check_hip(N.hip().bw_hip_stream_release(dev._h, C.c_void_p(s2.cuda_stream)))


def test_noalloc_rejects(ops, gpu, monkeypatch):
    import torch
    host, dev = ops[0]
    n = host.size
    x = torch.zeros(n, dtype=torch.float64, device=gpu)
    y = torch.zeros(n, dtype=torch.float64, device=gpu)
    # an operator whose scratch exceeds the array: the limit lowered for the test
    monkeypatch.setenv("BWHIP_NOALLOC_LIMIT", "4096")
    rc = _noalloc(dev, x, y, N.BW_ORDER_ORIGINAL, N.BW_PREC_FP64)
    assert rc == N.BW_ESHAPE
    assert "bytes of scratch" in N.hip().bw_hip_last_error().decode()
    monkeypatch.delenv("BWHIP_NOALLOC_LIMIT")
    assert _noalloc(dev, x, y, 7, N.BW_PREC_FP64) == N.BW_EINVAL
    assert _noalloc(dev, x, y, N.BW_ORDER_ORIGINAL, 3) == N.BW_EINVAL
    sel = HipHmat.selection(host, np.arange(10), np.arange(20), device=0)
    xs = torch.zeros(3 * 20, dtype=torch.float64, device=gpu)
    ys = torch.zeros(3 * 10, dtype=torch.float64, device=gpu)
    assert _noalloc(sel, xs, ys, N.BW_ORDER_ORIGINAL, N.BW_PREC_FP64) == N.BW_ENOTSUP
    sel.close()
    torch.cuda.synchronize()
