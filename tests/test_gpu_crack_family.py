"""The benchmark's 16 uneven square cracks (3DR0-H, element counts
log-uniform in [1 000, 40 000], seed 7: 477 402 DoF, E = 2.96e9) as one
block-diagonal operator (bw_hip_create_multi), against the oracle
(restatement of Hmat::matvecOriginal, src/hmat/hmatrix/hmat.cc:572-715, and of
its transpose) applied mesh by mesh: H x in FP64 at 1e-12 and FP32 at 1e-5,
H^T x, graph replays bit for bit equal to the eager call, and 32 right-hand
sides through one batched graph replay, every mesh slice and column."""
import time

import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HipHmat, HostHmat
from bigwham_dev_amd.meshes import crack_family, rect_grid

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def family(gpu):
    sides = crack_family(16, 1000, 40000, seed=7)
    t0 = time.perf_counter()
    hosts = []
    for s in sides:
        h = HostHmat.from_mesh("3DR0-H", *rect_grid(s), [1.0, 0.25])
        h.build(64, 3.0, 1e-6)
        hosts.append(h)
    print(f"crack family: sides {sides}, DoF {sum(h.size for h in hosts)}, "
          f"E {sum(h.nb_entries() for h in hosts)}, build {time.perf_counter() - t0:.1f}s",
          flush=True)
    dev = HipHmat.many(hosts, precisions=("fp64", "fp32"), adjoint=True, device=0)
    yield hosts, dev
    dev.close()


def test_family_matvec_adjoint(family, gpu):
    import torch
    hosts, dev = family
    off = dev.row_offsets
    assert list(np.diff(off)) == [h.size for h in hosts]
    assert dev.nb_entries == sum(O.nb_entries(h.flat()) for h in hosts)
    rng = np.random.default_rng(1234)
    x = rng.uniform(-1, 1, int(off[-1]))
    xd = torch.from_numpy(x).to(gpu)
    print("family: H x fp64 / fp32, H^T x", flush=True)
    y64 = dev.matvec(xd).cpu().numpy()
    y32 = dev.matvec(xd, precision="fp32").cpu().numpy()
    yt = dev.rmatvec(xd).cpu().numpy()
    yt32 = dev.rmatvec(xd, precision="fp32").cpu().numpy()
    for m, h in enumerate(hosts):
        sl = slice(off[m], off[m + 1])
        ref = O.matvec_original(h.flat(), x[sl])
        assert rel(y64[sl], ref) <= TOL64, m
        assert rel(y32[sl], ref) <= TOL32, m
        reft = O.rmatvec_original(h.flat(), x[sl])
        assert rel(yt[sl], reft) <= TOL64, m
        assert rel(yt32[sl], reft) <= TOL32, m
    # the identity <H x, w> = <x, H^T w> on the whole composite
    w = rng.uniform(-1, 1, x.size)
    lhs = float(np.dot(y64, w))
    rhs = float(np.dot(x, dev.rmatvec(torch.from_numpy(w).to(gpu)).cpu().numpy()))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_family_graph_replay(family, gpu, precision):
    import torch
    hosts, dev = family
    n = int(dev.row_offsets[-1])
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.uniform(-1, 1, n)).to(gpu)
    print(f"family: graph replay {precision}", flush=True)
    g = dev.graph(x, precision=precision)
    for rep in range(3):
        x.copy_(torch.from_numpy(rng.uniform(-1, 1, n)).to(gpu))
        eager = dev.matvec(x, precision=precision).clone()
        torch.cuda.synchronize()
        assert torch.equal(g.launch(), eager), rep
        torch.cuda.synchronize()
    g.close()


def test_family_batched_graph(family, gpu):
    """32 right-hand sides over all 16 meshes in one graph replay (FP32 and
    FP64 MFMA), every mesh slice and every column against the oracle."""
    import torch
    hosts, dev = family
    off = dev.row_offsets
    n = int(off[-1])
    X = np.random.default_rng(77).uniform(-1, 1, (n, 32))
    Xd = torch.from_numpy(X).to(gpu)
    refs = []
    for m, h in enumerate(hosts):
        sl = slice(off[m], off[m + 1])
        f = h.flat()
        refs.append(np.stack([O.matvec_original(f, X[sl, j]) for j in range(32)], 1))
    for precision, tol in (("fp32", TOL32), ("fp64", TOL64)):
        print(f"family: 32-RHS batched graph {precision}", flush=True)
        g = dev.graph_batched(Xd, precision=precision)
        Y = g.launch()
        torch.cuda.synchronize()
        eager = dev.matvec_batched(Xd, precision=precision)
        assert torch.equal(Y, eager)
        Yh = Y.cpu().numpy()
        for m in range(len(hosts)):
            sl = slice(off[m], off[m + 1])
            for j in range(32):
                assert rel(Yh[sl, j], refs[m][:, j]) <= tol, (precision, m, j)
        g.close()
