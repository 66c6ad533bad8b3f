"""The host library of bigwham_dev_amd against fixtures dumped from the compiled
reference harness (tests/golden/make_ref_fixtures.py, oracle/_ref built from
the reference's clustering, pattern and segment-kernel sources), so the pins
hold where the reference sources are not mounted.  CPU only."""
from pathlib import Path

import numpy as np
import pytest

from bigwham_dev_amd import HostHmat
from bigwham_dev_amd import _native as N
from bigwham_dev_amd.meshes import penny_delaunay, rect_grid, segment_line

FIX = Path(__file__).resolve().parent / "golden" / "ref_pattern_fixtures.npz"
MESHES = {"2DS0-H": (lambda: segment_line(2000), 32),
          "3DR0-H": (lambda: rect_grid(60), 64),
          "3DT0-H": (lambda: penny_delaunay(3000, seed=5), 64)}


@pytest.fixture(scope="module")
def fix():
    return np.load(FIX, allow_pickle=False)


@pytest.mark.parametrize("kernel", list(MESHES))
@pytest.mark.parametrize("homogeneous", [0, 1])
def test_pattern_and_permutation(fix, kernel, homogeneous):
    """Cluster tree (exchange-sort order), permutation and block pattern are
    the reference's own (cluster.cc:376-718, h_pattern.cc:87-166)."""
    mesh, leaf = MESHES[kernel]
    coor, conn = mesh()
    h = HostHmat.from_mesh(kernel, coor, conn, [1.0, 0.25],
                           homogeneous_size_pattern=bool(homogeneous))
    h.build_pattern(leaf, 3.0)
    key = f"{kernel}_h{homogeneous}"
    assert np.array_equal(h.collocation_points(), fix[key + "_points"])
    assert np.array_equal(h.permutation(), fix[key + "_perm"])
    assert np.array_equal(h.fr_pattern()[:5], fix[key + "_fr"])
    assert np.array_equal(h.lr_pattern()[:5], fix[key + "_lr"])


def test_segment_kernel(fix):
    """We_segment_0 (2DS0 influence of a uniform-DD segment) on 20 receivers,
    seen through the influence of a two-element mesh whose source segment
    lies on the x axis: traction = receiver frame applied to the stress."""
    rng = np.random.default_rng(11)
    for (size, G, nu, x0, x1), W in zip(fix["we_segment_0_args"], fix["we_segment_0_values"]):
        E = 2.0 * G * (1.0 + nu)
        th = rng.uniform(0, np.pi)
        t = np.array([np.cos(th), np.sin(th)])
        p = np.array([[-size / 2, 0.0], [size / 2, 0.0], [0.0, 0.0], [0.0, 0.0]])
        h0 = HostHmat.from_mesh("2DS0-H", p[:2], np.array([[0, 1]]), [E, nu])
        R0 = _rotations(h0, 1)[0]
        c = R0.T @ np.array([x0, x1])  # receiver centre with local coordinates (x0, x1)
        p[2], p[3] = c - 0.05 * t, c + 0.05 * t
        h = HostHmat.from_mesh("2DS0-H", p, np.array([[0, 1], [2, 3]]), [E, nu])
        R = _rotations(h, 2)
        xl = R[0] @ (p[2:].mean(0) - p[:2].mean(0))
        assert np.allclose(xl, [x0, x1], atol=1e-14)
        W = W.reshape(2, 3)
        n = R[0] @ R[1][1]
        s = R[0] @ R[1][0]
        exp = np.array([[n[0] * s[0] * W[0, 0] + (n[0] * s[1] + n[1] * s[0]) * W[0, 1] + n[1] * s[1] * W[0, 2],
                         n[0] * s[0] * W[1, 0] + (n[0] * s[1] + n[1] * s[0]) * W[1, 1] + n[1] * s[1] * W[1, 2]],
                        [n[0] ** 2 * W[0, 0] + 2 * n[0] * n[1] * W[0, 1] + n[1] ** 2 * W[0, 2],
                         n[0] ** 2 * W[1, 0] + 2 * n[0] * n[1] * W[1, 1] + n[1] ** 2 * W[1, 2]]])
        got = h.influence(0, 1)
        assert np.allclose(got, exp, rtol=1e-12, atol=1e-12 * np.abs(exp).max())


def _rotations(h, n_elts):
    out = np.empty(n_elts * 4)
    N.check_host(N.host().bwh_rotation_matrices(h._h, N.dptr(out)))
    return out.reshape(n_elts, 2, 2)


def test_penny_105k_pattern():
    """The 600k-DoF penny-crack mesh (meshes.penny_delaunay(105000, seed=0)):
    clusters of more than 60 000 points there have exactly tied split
    coordinates, where the exchange-sort replay stops after a swap budget
    (hb_cluster.cpp).  The permutation and the block pattern still equal the
    compiled reference's (tests/golden/make_c3_pattern_fixture.py)."""
    import hashlib
    f = np.load(Path(__file__).resolve().parent / "golden" / "ref_pattern_penny_105k.npz",
                allow_pickle=False)
    digest = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    coor, conn = penny_delaunay(105000, seed=0)
    h = HostHmat.from_mesh("3DT0-H", coor, conn, [1.0, 0.25])
    assert digest(h.collocation_points()) == str(f["points_sha256"]), "mesh differs from the fixture's"
    h.build_pattern(64, 3.0)
    assert np.array_equal(h.permutation(), f["perm"].astype(np.int64))
    fr, lr = h.fr_pattern(), h.lr_pattern()
    assert (fr.shape[1], lr.shape[1]) == (int(f["n_fr"]), int(f["n_lr"]))
    assert digest(fr[:5].astype(np.int64)) == str(f["fr_sha256"])
    assert digest(lr[:5].astype(np.int64)) == str(f["lr_sha256"])
