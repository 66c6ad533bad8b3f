"""Edge cases of the device path against the oracle (restatement of
Hmat::matvec / matvecOriginal, src/hmat/hmatrix/hmat.cc:472-715): a single
element, matrices with no low-rank block, fixed-rank and homogeneous-size
patterns (hmat_cuda.cc's only supported layout, bigwham_io.cc:89-96), a
single right-hand side through the batched call, a one-point selection and a
one-matrix composite."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HipHmat, HostHmat
from bigwham_dev_amd.meshes import rect_grid, segment_line

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _check(h, gpu, seed=0):
    import torch
    dev = HipHmat(h, precisions=("fp64", "fp32"), adjoint=True)
    x = np.random.default_rng(seed).uniform(-1, 1, h.size)
    xd = torch.from_numpy(x).to(gpu)
    for original in (False, True):
        ref = (O.matvec_original if original else O.matvec)(h.flat(), x)
        assert rel(dev.matvec(xd, original=original).cpu().numpy(), ref) <= 1e-12
        assert rel(dev.matvec(xd, original=original, precision="fp32").cpu().numpy(), ref) <= 1e-5
        reft = (O.rmatvec_original if original else O.rmatvec)(h.flat(), x)
        assert rel(dev.rmatvec(xd, original=original).cpu().numpy(), reft) <= 1e-12
    return dev


def test_single_element(gpu):
    coor, conn = segment_line(1)
    h = HostHmat.from_mesh("2DS0-H", coor, conn, [1.0, 0.25])
    h.build(32, 3.0, 1e-6)
    assert h.size == 2 and h.flat().n_lr == 0
    print("single element", flush=True)
    _check(h, gpu)


def test_full_rank_only(gpu):
    coor, conn = rect_grid(5)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    assert h.flat().n_lr == 0
    print("full rank only", flush=True)
    _check(h, gpu)


@pytest.mark.parametrize("fixed_rank,homogeneous", [(3, False), (-1, True), (4, True)])
def test_pattern_variants(gpu, fixed_rank, homogeneous):
    coor, conn = rect_grid(22)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25], fixed_rank=fixed_rank,
                           homogeneous_size_pattern=homogeneous)
    h.build(32, 3.0, 1e-6)
    print(f"pattern fixed_rank={fixed_rank} homogeneous={homogeneous}", flush=True)
    _check(h, gpu, seed=3)


def test_batched_single_column_and_tiny_selection(gpu):
    import torch
    coor, conn = rect_grid(16)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    dev = HipHmat(h, precisions=("fp64", "fp32"))
    X = torch.from_numpy(np.random.default_rng(5).uniform(-1, 1, (h.size, 1))).to(gpu)
    ref = O.matvec_original(h.flat(), X[:, 0].cpu().numpy())
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        Y = dev.matvec_batched(X, precision=prec)
        assert Y.shape == (h.size, 1)
        assert rel(Y[:, 0].cpu().numpy(), ref) <= tol
    s = HipHmat.selection(h, [7], [7])
    xs = np.array([1.0, -2.0, 0.5])
    ys = s.matvec(xs)
    assert rel(ys, O.selection_matvec(h.flat(), [7], [7], xs)) <= 1e-12
    one = HipHmat.many([h], precisions=("fp64",))
    xd = X[:, 0].contiguous()
    assert rel(one.matvec(xd).cpu().numpy(), ref) <= 1e-12


def test_wrong_lengths_are_refused(gpu):
    """Vectors of the wrong length or type never reach a kernel."""
    import torch
    coor, conn = rect_grid(6)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    dev = HipHmat(h, precisions=("fp64",))
    n = h.size
    print(f"wrong lengths n={n}", flush=True)
    for bad in (np.zeros(n - 1), np.zeros(n + 3)):
        with pytest.raises(ValueError):
            dev.matvec(bad)
        with pytest.raises(ValueError):
            dev.matvec(torch.from_numpy(bad).to(gpu))
    with pytest.raises(ValueError):
        dev.matvec(np.zeros(n), out=np.zeros(n - 1))
    with pytest.raises(ValueError):
        dev.matvec(torch.zeros(n, dtype=torch.float64, device=gpu),
                   out=torch.zeros(n, dtype=torch.float32, device=gpu))
    with pytest.raises(ValueError):
        dev.graph(torch.zeros(n - 1, dtype=torch.float64, device=gpu))
    with pytest.raises(ValueError):
        dev.matvec_batched(torch.zeros((n - 1, 4), dtype=torch.float64, device=gpu))
    # and the right lengths still work afterwards
    x = np.random.default_rng(3).uniform(-1, 1, n)
    assert rel(dev.matvec(x), O.matvec_original(h.flat(), x)) <= 1e-12
