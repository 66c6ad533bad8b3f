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
