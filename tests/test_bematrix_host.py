"""BEMatrix helpers that need no GPU (pattern-only construction,
directly_build=False): element frames and the local/global conversions of
Mesh::ConvertToGlobal / ConvertToLocal (src/core/mesh.h:91-125)."""
import numpy as np

from bigwham_dev_amd.bigwham4py import BEMatrix
from bigwham_dev_amd.meshes import penny_delaunay, rect_grid, segment_line


def _pattern_only(kernel, mesh):
    coor, conn = mesh
    return BEMatrix(kernel, coor, conn, [1.0, 0.25], max_leaf_size=32, directly_build=False,
                    verbose=False)


def test_frames_2d():
    M = _pattern_only("2DS0-H", segment_line(10))
    n = M.get_element_normals().reshape(-1, 2)
    R = M.get_rotation_matrix().reshape(-1, 2, 2)
    assert n.shape == (10, 2) and R.shape == (10, 2, 2)
    for Re in R:  # orthonormal
        assert np.allclose(Re @ Re.T, np.eye(2), atol=1e-14)
    # a crack along x: the normal is +-y, and local normal component of a
    # global vector along that normal is +-1 (local = (shear, normal))
    assert np.allclose(np.abs(n[:, 1]), 1.0) and np.allclose(n[:, 0], 0.0, atol=1e-14)
    g = np.tile(n[0], 10)
    loc = M.convert_to_local(g).reshape(-1, 2)
    assert np.allclose(np.abs(loc[:, 1]), 1.0) and np.allclose(loc[:, 0], 0.0, atol=1e-14)


def test_round_trip_3d():
    for kernel, mesh in (("3DR0-H", rect_grid(5)), ("3DT0-H", penny_delaunay(200, seed=2))):
        M = _pattern_only(kernel, mesh)
        R = M.get_rotation_matrix().reshape(-1, 3, 3)
        x = np.random.default_rng(1).uniform(-1, 1, 3 * R.shape[0])
        xl = M.convert_to_local(x)
        assert np.allclose(M.convert_to_global(xl), x, atol=1e-14)
        # the local normal component is the projection on the element normal
        nrm = M.get_element_normals().reshape(-1, 3)
        assert np.allclose(xl.reshape(-1, 3)[:, 2],
                           np.einsum("ei,ei->e", nrm, x.reshape(-1, 3)), atol=1e-14)


def test_plot_pattern():
    """One patch per block of the pattern, inside the square of the points."""
    import matplotlib
    matplotlib.use("Agg")
    M = _pattern_only("3DR0-H", rect_grid(8))
    fig = M.plotPattern(plot_index=True)
    pat = M._getPattern()
    coll = fig.axes[0].collections[0]
    assert len(coll.get_paths()) == pat.shape[0]
    assert fig.axes[0].get_xlim() == (0, pat[:, 2].max())
