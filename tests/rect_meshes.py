"""Source / receiver mesh pairs for the rectangular operators of the tests:
each receiver mesh is a different mesh of the same element type (another
crack, offset, tilted or scaled), so no receiver point sits on a source
element's edge."""
import numpy as np

from bigwham_dev_amd.meshes import penny_delaunay, rect_grid, segment_line


def _rot_x(coor, deg):
    a = np.deg2rad(deg)
    R = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    return coor @ R.T


def pair(kernel: str, small: bool = False):
    """(coor_src, conn_src, coor_rec, conn_rec, leaf) for one rectangular
    kernel: about 2 000-6 000 source elements, or a few hundred (leaf 8)
    when small, for dense checks."""
    if kernel == "2DS0-2DS0-H":
        cs, ks = segment_line(300 if small else 2000)
        cr, kr = segment_line(190 if small else 1300, L=1.3)
        cr = cr + np.array([0.0, 0.05])
        return cs, ks, cr, kr, 8 if small else 32
    if kernel == "3DR0-3DR0-H":
        cs, ks = rect_grid(14 if small else 60)
        cr, kr = rect_grid(*((11, 9) if small else (45, 38)), half=0.8)
        cr = _rot_x(cr, 20.0) + np.array([0.11, -0.07, 0.3])
        return cs, ks, cr, kr, 8 if small else 64
    if kernel == "3DT0-3DT0-H":
        cs, ks = penny_delaunay(150 if small else 3000, seed=5)
        cr, kr = penny_delaunay(100 if small else 1800, seed=9)
        cr = _rot_x(0.7 * cr, -15.0) + np.array([0.2, 0.0, 0.25])
        return cs, ks, cr, kr, 8 if small else 64
    if kernel == "3DR0-3DR0-H-mode1":
        cs, ks = rect_grid(16 if small else 40)
        cr, kr = rect_grid(12 if small else 30)
        cr = cr + np.array([0.5123, 0.3141, 0.0])
        return cs, ks, cr, kr, 8 if small else 64
    raise ValueError(kernel)


KERNELS = ["2DS0-2DS0-H", "3DR0-3DR0-H", "3DT0-3DT0-H", "3DR0-3DR0-H-mode1"]


def edge_pairs():
    """Degenerate sides: one source element against a grid of receivers and
    a grid of sources against one receiver (kernel, coor_src, conn_src,
    coor_rec, conn_rec, leaf)."""
    one, k1 = rect_grid(1, half=0.3)
    grid, kg = rect_grid(10)
    lifted = grid + np.array([0.0, 0.0, 0.2])
    one_up = one + np.array([0.05, -0.02, 0.15])
    s1, q1 = segment_line(1, L=0.2)
    sl, ql = segment_line(50)
    return {"one_source": ("3DR0-3DR0-H", one, k1, lifted, kg, 8),
            "one_receiver": ("3DR0-3DR0-H", grid, kg, one_up, k1, 8),
            "segment_one_source": ("2DS0-2DS0-H", s1, q1, sl + np.array([0.0, 0.1]), ql, 8),
            "segment_one_receiver": ("2DS0-2DS0-H", sl, ql, s1 + np.array([0.03, 0.1]), q1, 8)}
