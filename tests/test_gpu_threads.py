"""Host threads calling the engine at once (the C-ABI's threading contract):
several threads on one handle and one stream, two handles driven from two
threads on two streams of their own, and one handle driven on two streams at
once (each stream has its own device scratch).  Every result must equal the
single-threaded one bit for bit."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from bigwham_dev_amd import HipHmat, HostHmat
from bigwham_dev_amd.meshes import rect_grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def host():
    coor, conn = rect_grid(30)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    return h


def test_threads_share_one_handle(gpu, host):
    import torch
    dev = HipHmat(host, precisions=("fp64", "fp32"))
    xs = [torch.from_numpy(np.random.default_rng(k).uniform(-1, 1, host.size)).to(gpu)
          for k in range(8)]
    ref = [dev.matvec(x).cpu().numpy() for x in xs]
    ref32 = [dev.matvec(x, precision="fp32").cpu().numpy() for x in xs]
    torch.cuda.synchronize()
    print(f"threads on one handle n={host.size}", flush=True)

    def work(k):
        out = []
        for _ in range(10):  # same (default) stream: the calls are ordered on the device
            out.append((dev.matvec(xs[k]), dev.matvec(xs[k], precision="fp32")))
        torch.cuda.synchronize()
        return k, [(a.cpu().numpy(), b.cpu().numpy()) for a, b in out]

    with ThreadPoolExecutor(4) as ex:
        for k, res in ex.map(work, range(8)):
            for y, y32 in res:
                assert np.array_equal(y, ref[k]) and np.array_equal(y32, ref32[k])


def test_two_handles_two_streams(gpu, host):
    import torch
    devs = [HipHmat(host, precisions=("fp64",)) for _ in range(2)]
    x = torch.from_numpy(np.random.default_rng(3).uniform(-1, 1, host.size)).to(gpu)
    ref = devs[0].matvec(x).cpu().numpy()
    print("two handles, two streams", flush=True)

    def work(i):
        s = torch.cuda.Stream(device=gpu)  # default priority
        s.wait_stream(torch.cuda.default_stream(gpu))  # x was written there
        with torch.cuda.stream(s):
            x_i = x.clone()
            ys = [devs[i].matvec(x_i) for _ in range(20)]
        s.synchronize()
        return [y.cpu().numpy() for y in ys]

    with ThreadPoolExecutor(2) as ex:
        for res in ex.map(work, range(2)):
            for y in res:
                assert np.array_equal(y, ref)


def test_one_handle_two_streams(gpu, host):
    """One handle, two streams, overlapping on the device: single-vector
    matvecs in both precisions, H^T and the batched product, each stream with
    its own inputs.  Results equal the serial run bit for bit."""
    import torch
    dev = HipHmat(host, precisions=("fp64", "fp32"), adjoint=True)
    n = host.size
    rng = np.random.default_rng(17)
    xs = [torch.from_numpy(rng.uniform(-1, 1, n)).to(gpu) for _ in range(2)]
    Xs = [torch.from_numpy(rng.uniform(-1, 1, (n, 32))).to(gpu) for _ in range(2)]

    def calls(i):
        return (dev.matvec(xs[i]), dev.matvec(xs[i], precision="fp32"), dev.rmatvec(xs[i]),
                dev.matvec_batched(Xs[i], precision="fp32"), dev.matvec_batched(Xs[i]))

    serial = [[t.cpu().numpy() for t in calls(i)] for i in range(2)]
    torch.cuda.synchronize()
    print("one handle, two streams", flush=True)
    streams = [torch.cuda.Stream(device=gpu) for _ in range(2)]  # default priority
    for s in streams:
        s.wait_stream(torch.cuda.default_stream(gpu))

    def work(i):
        outs = []
        with torch.cuda.stream(streams[i]):
            for _ in range(6):  # long enough on the device for the streams to overlap
                outs.append(calls(i))
        streams[i].synchronize()
        return [[t.cpu().numpy() for t in o] for o in outs]

    with ThreadPoolExecutor(2) as ex:
        res = list(ex.map(work, range(2)))
    for i in range(2):
        for o in res[i]:
            for got, want in zip(o, serial[i]):
                assert np.array_equal(got, want), i
    for s in streams:
        dev.release_stream(s.cuda_stream)
    # a released stream allocates again on its next call
    with torch.cuda.stream(streams[0]):
        y = dev.matvec(xs[0])
    streams[0].synchronize()
    assert np.array_equal(y.cpu().numpy(), serial[0][0])
