"""Sub-matrix selection on the GPU (bw_hip_create_selection) against the
oracle's restatement of HmatSelection::matvecOriginal / matvec
(src/hmat/hmatrix/hmat_selection.cc:188-257)."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import BigWhamError, HipHmat, HostHmat
from bigwham_dev_amd.meshes import rect_grid

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def host():
    coor, conn = rect_grid(24)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    return h


@pytest.mark.parametrize("original", [True, False])
def test_selection_matvec(gpu, host, original):
    import torch
    d = host.dof
    npts = host.size // d
    rng = np.random.default_rng(12)
    rows = rng.choice(npts, 150, replace=False)
    cols = np.concatenate([rng.choice(npts, 200, replace=False), [7, 7]])  # repeated point
    print(f"selection original={original} rows={rows.size} cols={cols.size}", flush=True)
    dev = HipHmat.selection(host, rows, cols, original=original, precisions=("fp64", "fp32"),
                            adjoint=True)
    assert (dev.n_rows, dev.n_cols) == (d * rows.size, d * cols.size)
    f = host.flat()
    x = rng.uniform(-1, 1, d * cols.size)
    ref = O.selection_matvec(f, rows, cols, x, original=original)
    xd = torch.from_numpy(x).to(gpu)
    y = dev.matvec(xd, original=original).cpu().numpy()
    assert rel(y, ref) <= 1e-12
    assert rel(dev.matvec(xd, original=original, precision="fp32").cpu().numpy(), ref) <= 1e-5
    assert rel(dev.matvec(x, original=original), ref) <= 1e-12  # host pointers
    xr = rng.uniform(-1, 1, d * rows.size)
    reft = O.selection_matvec(f, rows, cols, xr, original=original, adjoint=True)
    yt = dev.rmatvec(torch.from_numpy(xr).to(gpu), original=original).cpu().numpy()
    assert rel(yt, reft) <= 1e-12
    # only the blocks meeting the selection are stored: a cluster of points
    # touches a fraction of the blocks (random points touch them all)
    near = host.permutation()[:20] if original else np.arange(20)
    small = HipHmat.selection(host, near, near, original=original)
    assert small.nb_entries < host.nb_entries() // 4
    xs = rng.uniform(-1, 1, d * near.size)
    ys = small.matvec(torch.from_numpy(xs).to(gpu), original=original).cpu().numpy()
    assert rel(ys, O.selection_matvec(f, near, near, xs, original=original)) <= 1e-12
    with pytest.raises(BigWhamError, match="other ordering"):
        dev.matvec(xd, original=not original)
    with pytest.raises(BigWhamError, match="selection"):
        dev.diagonal()
    g = dev.graph(xd, original=original)
    for _ in range(3):
        assert torch.equal(g.launch().clone(), dev.matvec(xd, original=original))
    g.close()


def test_selection_errors(gpu, host):
    with pytest.raises(BigWhamError, match="out of bounds"):
        HipHmat.selection(host, [0, host.size], [1])
    with pytest.raises(BigWhamError, match="empty"):
        HipHmat.selection(host, [], [1])
