"""Rectangular operators built on the host (bwh_create_rect): a source mesh
(columns) against the collocation points of another mesh (rows), as the
reference's BigWhamIO rectangular constructor and
HRepresentationRectangularMatrix build them (bigwham_io.cc:292-484,
hierarchical_representation.h:88-137).  CPU only.

Pins:
- cluster permutations of both sides and the I x J block pattern equal the
  compiled reference's (oracle/_ref, cluster.cc:213-374) and the fixtures
  dumped from it (tests/golden/ref_pattern_rect_fixtures.npz);
- every entry of the rectangular operator equals the receivers x sources
  block of the square operator on the union of the two meshes, so rows use
  the receiver elements and their frames and columns the source elements;
- the H-matrix (oracle restatement of Hmat::matvecOriginal / its adjoint)
  stays within the ACA tolerance of that dense operator."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import BigWhamError, HostHmat
from layout_emu import LayoutDump
from rect_meshes import KERNELS, edge_pairs, pair

ROOT = Path(__file__).resolve().parents[1]
REF = ROOT / "oracle" / "_ref" / "libbwref.so"
FIX = ROOT / "tests" / "golden" / "ref_pattern_rect_fixtures.npz"
SQUARE = {"2DS0-2DS0-H": "2DS0-H", "3DR0-3DR0-H": "3DR0-H", "3DT0-3DT0-H": "3DT0-H",
          "3DR0-3DR0-H-mode1": "3DR0-H-mode1"}


def _host(kernel, hom=False, small=False):
    cs, ks, cr, kr, leaf = pair(kernel, small)
    h = HostHmat.from_meshes(kernel, cs, ks, cr, kr, [1.0, 0.25], homogeneous_size_pattern=hom)
    return h, leaf, (cs, ks, cr, kr)


@pytest.fixture(scope="module")
def fix():
    return np.load(FIX, allow_pickle=False)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("hom", [0, 1])
def test_pattern_matches_reference_fixtures(fix, kernel, hom):
    h, leaf, _ = _host(kernel, bool(hom))
    h.build_pattern(leaf, 3.0)
    assert np.array_equal(h.receiver_collocation_points(), fix[kernel + "_rec_points"])
    assert np.array_equal(h.collocation_points(), fix[kernel + "_src_points"])
    key = f"{kernel}_h{hom}"
    assert np.array_equal(h.permutation_receivers(), fix[key + "_perm0"])
    assert np.array_equal(h.permutation(), fix[key + "_perm1"])
    assert np.array_equal(h.fr_pattern()[:5], fix[key + "_fr"])
    assert np.array_equal(h.lr_pattern()[:5], fix[key + "_lr"])
    f = h.flat()
    assert (f.n_row_pts, f.n_col_pts) == h.shape_points()
    assert h.shape == (h.dof * f.n_row_pts, h.dof * f.n_col_pts)


@pytest.mark.skipif(not REF.exists(), reason="reference harness oracle/_ref not built")
@pytest.mark.parametrize("hom", [0, 1])
def test_pattern_matches_compiled_reference(hom):
    from golden.make_ref_rect_fixtures import reference_pattern
    ref = C.CDLL(str(REF))
    for kernel in ["2DS0-2DS0-H", "3DR0-3DR0-H"]:
        h, leaf, _ = _host(kernel, bool(hom))
        h.build_pattern(leaf, 2.0)  # another eta than the fixtures'
        q0, q1, fr, lr = reference_pattern(ref, h.receiver_collocation_points(),
                                           h.collocation_points(), leaf, 2.0, hom)
        assert np.array_equal(h.permutation_receivers(), q0)
        assert np.array_equal(h.permutation(), q1)
        assert np.array_equal(h.fr_pattern()[:5], fr)
        assert np.array_equal(h.lr_pattern()[:5], lr)


@pytest.mark.parametrize("kernel", KERNELS)
def test_entries_are_the_union_mesh_block(kernel):
    h, leaf, (cs, ks, cr, kr) = _host(kernel, small=True)
    h.build_pattern(leaf, 3.0)
    u = HostHmat.from_mesh(SQUARE[kernel], np.vstack([cs, cr]), np.vstack([ks, kr + len(cs)]),
                           [1.0, 0.25])
    u.build_pattern(leaf, 3.0)
    d, ns = h.dof, len(ks)
    D = h.dense(original=True)
    assert D.shape == (d * len(kr), d * ns)
    assert np.array_equal(D, u.dense(original=True)[d * ns:, : d * ns])
    assert np.all(np.isfinite(D))
    # one entry through the kernel helper: receiver element 3, source element 5
    assert np.array_equal(h.influence(5, 3), D[d * 3:d * 4, d * 5:d * 6])


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("hom", [False, True])
def test_hmat_within_aca_tolerance(kernel, hom):
    h, leaf, _ = _host(kernel, hom, small=True)
    h.build(leaf, 3.0, 1e-6)
    f = h.flat()
    assert f.n_lr > 0  # (the two parallel segment lines are all low rank without hom)
    D = h.dense(original=True)
    Dp = h.dense(original=False)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, D.shape[1])
    u = rng.uniform(-1, 1, D.shape[0])

    def rel(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)

    assert rel(O.matvec_original(f, x), D @ x) < 2e-5
    assert rel(O.rmatvec_original(f, u), D.T @ u) < 2e-5
    assert rel(O.matvec(f, x), Dp @ x) < 2e-5
    # the two orderings are the same operator under perm0 (rows) / perm1
    # (columns); one thread, so both sum in the same order
    d = h.dof
    ar = np.arange(d)[None, :]
    ri = (d * h.permutation_receivers()[:, None] + ar).ravel()
    ci = (d * h.permutation()[:, None] + ar).ravel()
    assert np.array_equal(O.matvec_original(f, x, n_threads=1)[ri],
                          O.matvec(f, x[ci], n_threads=1))
    assert O.nb_entries(f) == h.nb_entries()
    assert h.compression_ratio() == h.nb_entries() / (D.shape[0] * D.shape[1])
    assert h.max_aca_error() >= 0.0


def test_bad_inputs_are_refused():
    cs, ks, cr, kr, _ = pair("3DR0-3DR0-H", small=True)
    with pytest.raises(BigWhamError, match="unknown kernel"):
        HostHmat.from_meshes("3DR0-H", cs, ks, cr, kr, [1.0, 0.25])
    bad = kr.copy()
    bad[0, 0] = len(cr)
    with pytest.raises(BigWhamError, match="receiver mesh"):
        HostHmat.from_meshes("3DR0-3DR0-H", cs, ks, cr, bad, [1.0, 0.25])
    bad = ks.copy()
    bad[-1, 2] = -1
    with pytest.raises(BigWhamError, match="source mesh"):
        HostHmat.from_meshes("3DR0-3DR0-H", cs, bad, cr, kr, [1.0, 0.25])


def test_python_class_before_build():
    """BEMatrixRectangular (bigwham4py.py:491-565) with directly_build=False:
    the pattern and both permutations, no device work."""
    from bigwham_dev_amd.bigwham4py import BEMatrixRectangular
    cs, ks, cr, kr, leaf = pair("3DT0-3DT0-H", small=True)
    M = BEMatrixRectangular("3DT0-3DT0-H", cs, ks, cr, kr, np.array([1.0, 0.25]),
                            max_leaf_size=leaf, directly_build=False, verbose=False)
    assert M.shape == (3 * len(kr), 3 * len(ks))
    assert M.getSpatialDimension() == 3
    assert np.array_equal(M.getMeshCollocationPoints(), M.H_.receiver_collocation_points())
    assert sorted(M.getPermutationReceivers()) == list(range(len(kr)))
    assert sorted(M.getPermutation()) == list(range(len(ks)))
    pat = M._getPattern()
    assert len(pat) == M.H_.flat().n_fr + M.H_.flat().n_lr
    with pytest.raises(ValueError, match="Invalid kernel"):
        BEMatrixRectangular("3DT0-H", cs, ks, cr, kr, np.array([1.0, 0.25]),
                            directly_build=False)


def test_store_round_trip(tmp_path):
    """A rectangular operator saved and loaded (bigwham_dev_amd.store): both
    permutations and every block come back, so the oracle's product is the
    same to the bit."""
    from bigwham_dev_amd import store
    h, leaf, _ = _host("3DR0-3DR0-H", True, small=True)
    h.build(leaf, 3.0, 1e-6)
    store.save(h, tmp_path / "rect.hmat")
    g = store.load(tmp_path / "rect.hmat")
    f, fg = h.flat(), g.flat()
    assert (fg.n_row_pts, fg.n_col_pts) == (f.n_row_pts, f.n_col_pts)
    x = np.random.default_rng(4).uniform(-1, 1, h.shape[1])
    assert np.array_equal(O.matvec_original(fg, x, n_threads=1),
                          O.matvec_original(f, x, n_threads=1))


@pytest.mark.parametrize("case", ["3DR0-3DR0-H", "3DT0-3DT0-H"] + list(edge_pairs()))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("adjoint", [False, True])
def test_layout_replay(case, precision, adjoint):
    """The device layout of a rectangular operator between two meshes, its
    degenerate sides included (one source or one receiver element): every
    address the kernels form stays in its buffer, and a numpy replay of the
    kernels equals the oracle in both numberings."""
    if case in KERNELS:
        cs, ks, cr, kr, leaf = pair(case, small=True)
        kernel = case
    else:
        kernel, cs, ks, cr, kr, leaf = edge_pairs()[case]
    h = HostHmat.from_meshes(kernel, cs, ks, cr, kr, [1.0, 0.25])
    h.build(leaf, 3.0, 1e-6)
    f = h.flat()
    lay = LayoutDump(f, precision, adjoint, 4096)
    lay.check_bounds()
    lay.check_xcd_cover()
    x = np.random.default_rng(6).uniform(-1, 1, lay.n_in)
    tol = 1e-13 if precision == "fp64" else 1e-6
    ref = (O.rmatvec if adjoint else O.matvec)(f, x)
    assert np.linalg.norm(lay.replay(x) - ref) <= tol * np.linalg.norm(ref)
    p0, p1 = h.permutation_receivers(), h.permutation()
    pin, pout = (p0, p1) if adjoint else (p1, p0)
    ref = (O.rmatvec_original if adjoint else O.matvec_original)(f, x)
    assert np.linalg.norm(lay.replay(x, pin, pout) - ref) <= tol * np.linalg.norm(ref)


def test_fixed_rank_and_homogeneous():
    """fixed_rank reaches ACA on a rectangular operator as on a square one:
    every low-rank block stops after fixed_rank pivots (dof * fixed_rank
    columns).  On this pair the size-conserving block tree makes every
    low-rank block leaf x leaf points (checked), so fixed_rank = leaf
    represents them exactly."""
    cs, ks, cr, kr, leaf = pair("3DR0-3DR0-H", small=True)
    for fixed in (4, leaf):
        h = HostHmat.from_meshes("3DR0-3DR0-H", cs, ks, cr, kr, [1.0, 0.25],
                                 homogeneous_size_pattern=True, fixed_rank=fixed)
        h.build(leaf, 3.0, 1e-6)
        lp = h.lr_pattern()
        assert lp.shape[1] > 0 and lp[5].max() <= 3 * fixed
        assert np.all(lp[3] - lp[1] == leaf) and np.all(lp[4] - lp[2] == leaf)
    D = h.dense(original=True)
    x = np.random.default_rng(2).uniform(-1, 1, D.shape[1])
    assert np.linalg.norm(O.matvec_original(h.flat(), x) - D @ x) <= 1e-12 * np.linalg.norm(D @ x)
