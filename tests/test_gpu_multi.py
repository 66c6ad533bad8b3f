"""Many uneven meshes in one operator (bw_hip_create_multi) and the
graph-captured matvec (bw_hip_graph_*), on the GPU, against the oracle
(restatement of Hmat::matvecOriginal, src/hmat/hmatrix/hmat.cc:572-715)
applied mesh by mesh."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HipHmat, HostHmat
from bigwham_dev_amd.meshes import rect_grid

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def meshes():
    out = []
    for n in (6, 12, 20, 9):
        coor, conn = rect_grid(n)
        h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
        h.build(64, 3.0, 1e-6)
        out.append(h)
    return out


def test_many_meshes_one_operator(gpu, meshes):
    import torch
    print(f"multi: sizes {[h.size for h in meshes]}", flush=True)
    dev = HipHmat.many(meshes, precisions=("fp64", "fp32"), adjoint=True)
    off = dev.row_offsets
    assert list(np.diff(off)) == [h.size for h in meshes]
    assert list(dev.col_offsets) == list(off)
    x = np.random.default_rng(21).uniform(-1, 1, int(off[-1]))
    xd = torch.from_numpy(x).to(gpu)
    y64 = dev.matvec(xd).cpu().numpy()
    y32 = dev.matvec(xd, precision="fp32").cpu().numpy()
    yt = dev.rmatvec(xd).cpu().numpy()
    for m, h in enumerate(meshes):
        sl = slice(off[m], off[m + 1])
        ref = O.matvec_original(h.flat(), x[sl])
        assert rel(y64[sl], ref) <= TOL64
        assert rel(y32[sl], ref) <= TOL32
        assert rel(yt[sl], O.rmatvec_original(h.flat(), x[sl])) <= TOL64
    d = dev.diagonal()
    for m, h in enumerate(meshes):
        sl = slice(off[m], off[m + 1])
        assert np.array_equal(d[sl], O.diagonal(h.flat(), original=True))


def test_many_meshes_recompressed_fp32(gpu, meshes):
    """The composite's FP32 copy with its low-rank blocks recompressed
    (fp32_rank_tol, as for one mesh): the blocks cover only the diagonal
    meshes, the FP32 products (H x, H^T x, 32 right-hand sides) stay within
    1e-5 of the oracle mesh by mesh, and the FP64 copy is untouched."""
    import torch
    dev = HipHmat.many(meshes, precisions=("fp64", "fp32"), adjoint=True, fp32_rank_tol=1e-7)
    ref_dev = HipHmat.many(meshes, precisions=("fp64", "fp32"))
    assert sum(dev.entries["fp32"][:3]) <= sum(ref_dev.entries["fp32"][:3])
    off = dev.row_offsets
    rng = np.random.default_rng(23)
    x = rng.uniform(-1, 1, int(off[-1]))
    xd = torch.from_numpy(x).to(gpu)
    y64 = dev.matvec(xd).cpu().numpy()
    y32 = dev.matvec(xd, precision="fp32").cpu().numpy()
    yt32 = dev.rmatvec(xd, precision="fp32").cpu().numpy()
    X = rng.uniform(-1, 1, (int(off[-1]), 32))
    Y32 = dev.matvec_batched(torch.from_numpy(X).to(gpu), precision="fp32").cpu().numpy()
    for m, h in enumerate(meshes):
        sl = slice(off[m], off[m + 1])
        ref = O.matvec_original(h.flat(), x[sl])
        assert rel(y64[sl], ref) <= TOL64
        assert rel(y32[sl], ref) <= TOL32
        assert rel(yt32[sl], O.rmatvec_original(h.flat(), x[sl])) <= TOL32
        for j in (0, 31):
            assert rel(Y32[sl, j], O.matvec_original(h.flat(), X[sl, j])) <= TOL32


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_match_eager(gpu, meshes, precision):
    """Replays of the captured matvec, with eager matvecs and torch work in
    between, equal the eager result bit for bit and read x at replay time."""
    import torch
    print(f"graph: precision {precision}", flush=True)
    dev = HipHmat.many(meshes, precisions=(precision,))
    n = int(dev.row_offsets[-1])
    rng = np.random.default_rng(5)
    x = torch.from_numpy(rng.uniform(-1, 1, n)).to(gpu)
    g = dev.graph(x, precision=precision)
    for rep in range(4):
        x.copy_(torch.from_numpy(rng.uniform(-1, 1, n)).to(gpu))
        eager = dev.matvec(x, precision=precision).clone()
        other = dev.matvec(torch.ones_like(x), precision=precision)  # eager work between
        torch.cuda.synchronize()
        y = g.launch()
        torch.cuda.synchronize()
        assert torch.equal(y, eager), f"replay {rep}"
        assert float(other.abs().sum()) > 0
    g.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("adjoint", [False, True])
def test_batched_graph_replays(gpu, meshes, precision, adjoint):
    """The batched product captured as one graph (bw_hip_graph_create_batched):
    37 right-hand sides (two 32-column rounds) over the many-mesh operator, X
    row-major and read at replay time; every replay equals the eager batched
    call bit for bit, with eager calls on the same handle between replays,
    and every mesh slice and column matches the oracle."""
    import torch
    print(f"batched graph: precision {precision} adjoint {adjoint}", flush=True)
    dev = HipHmat.many(meshes, precisions=(precision,), adjoint=adjoint)
    n = int(dev.row_offsets[-1])
    off = dev.row_offsets
    rng = np.random.default_rng(6)
    X = torch.from_numpy(rng.uniform(-1, 1, (n, 37))).to(gpu)
    g = dev.graph_batched(X, precision=precision, adjoint=adjoint)
    tol = TOL64 if precision == "fp64" else TOL32
    fn = O.rmatvec_original if adjoint else O.matvec_original
    for rep in range(3):
        X.copy_(torch.from_numpy(rng.uniform(-1, 1, (n, 37))).to(gpu))
        eager = dev.matvec_batched(X, precision=precision, adjoint=adjoint).clone()
        dev.matvec(torch.ones(n, dtype=torch.float64, device=gpu), precision=precision,
                   adjoint=adjoint)  # eager work on the handle between replays
        torch.cuda.synchronize()
        Y = g.launch()
        torch.cuda.synchronize()
        assert torch.equal(Y, eager), f"replay {rep}"
    Xh, Yh = X.cpu().numpy(), Y.cpu().numpy()
    for m, h in enumerate(meshes):
        sl = slice(off[m], off[m + 1])
        for j in range(37):
            assert rel(Yh[sl, j], fn(h.flat(), Xh[sl, j])) <= tol, (m, j)
    g.close()


def test_saved_operator_replays_bitwise(gpu, meshes, tmp_path):
    """An operator saved with bigwham_dev_amd.store and loaded back (memory
    mapped, no host build) gives the same device results bit for bit."""
    import torch
    from bigwham_dev_amd import store
    h = meshes[2]
    f = store.load(store.save(h, tmp_path / "h"))
    print(f"store: n={h.size}", flush=True)
    a = HipHmat(h, precisions=("fp64", "fp32"))
    b = HipHmat(f, precisions=("fp64", "fp32"))
    x = torch.from_numpy(np.random.default_rng(8).uniform(-1, 1, h.size)).to(gpu)
    for prec in ("fp64", "fp32"):
        assert torch.equal(a.matvec(x, precision=prec), b.matvec(x, precision=prec))
    many = HipHmat.many([f, meshes[0]], precisions=("fp64",))
    y = many.matvec(torch.cat([x, x[: meshes[0].size]]))
    ya = a.matvec(x)  # the composite splits its slabs differently: same sums, other order
    assert float(torch.linalg.vector_norm(y[: h.size] - ya) / torch.linalg.vector_norm(ya)) <= TOL64
