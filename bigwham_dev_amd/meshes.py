"""Synthetic crack meshes of the benchmark configurations (seeded, deterministic).

segment_line   straight 2-D crack of uniform segments (2DS0-H)
rect_grid      square planar crack of uniform rectangles (3DR0-H, 3DR0-H-mode1),
               rows of rectangles along x, stacked in y, as the reference's
               CartesianMesh examples (interfaces/python/examples/example_3DR0_kernel.py)
penny_delaunay penny-shaped crack of radius r: Delaunay triangulation of jittered
               points (3DT0-H)
crack_family   sides of n square cracks whose element counts are log-uniform
               in [lo, hi] (many uneven meshes in one operator)
"""
from __future__ import annotations

import numpy as np


def segment_line(n: int, L: float = 1.0):
    """n uniform segments on x in [-L, L], y = 0 (2DS0-H straight crack)."""
    xs = np.arange(n + 1) * (2.0 * L / n) - L
    coor = np.stack([xs, np.zeros_like(xs)], 1)
    conn = np.stack([np.arange(n), np.arange(1, n + 1)], 1)
    return coor, conn


def rect_grid(nx: int, ny: int | None = None, half: float = 1.0):
    """nx x ny uniform rectangles on [-half, half]^2, z = 0, row by row in y."""
    ny = nx if ny is None else ny
    gx = np.linspace(-half, half, nx + 1)
    gy = np.linspace(-half, half, ny + 1)
    X, Y = np.meshgrid(gx, gy)
    coor = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    a = (j * (nx + 1) + i).ravel()
    conn = np.stack([a, a + 1, a + nx + 2, a + nx + 1], 1)
    return coor, conn


def penny_delaunay(n_points: int, radius: float = 1.0, jitter: float = 0.2, seed: int = 0):
    """Triangulated disc of radius `radius` in z = 0: a hexagonal lattice of
    spacing h covering the disc with ~n_points nodes, each interior node moved
    by up to jitter*h (seeded), the boundary ring placed exactly on the circle,
    then scipy's Delaunay triangulation.  About 2*n_points triangles, oriented
    counter-clockwise (normal +z)."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    area = np.pi * radius * radius
    # hexagonal lattice density: 2 / (sqrt(3) h^2) points per unit area
    h = np.sqrt(2.0 * area / (np.sqrt(3.0) * n_points))
    nr = int(np.ceil(radius / h)) + 2
    ii, jj = np.meshgrid(np.arange(-nr, nr + 1), np.arange(-nr, nr + 1), indexing="ij")
    px = (ii + 0.5 * (jj % 2)) * h
    py = jj * h * np.sqrt(3.0) / 2.0
    pts = np.stack([px.ravel(), py.ravel()], 1)
    r = np.hypot(pts[:, 0], pts[:, 1])
    inner = pts[r < radius - 0.75 * h]
    inner = inner + rng.uniform(-jitter * h, jitter * h, inner.shape)
    nb = int(np.round(2.0 * np.pi * radius / h))
    th = np.arange(nb) * (2.0 * np.pi / nb)
    ring = np.stack([radius * np.cos(th), radius * np.sin(th)], 1)
    xy = np.concatenate([inner, ring], 0)
    tri = Delaunay(xy).simplices.astype(np.int64)
    # drop slivers on the boundary (all three nodes on the ring)
    on_ring = np.zeros(len(xy), dtype=bool)
    on_ring[len(inner):] = True
    tri = tri[~on_ring[tri].all(1)]
    a, b, c = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    flip = cross < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    coor = np.concatenate([xy, np.zeros((len(xy), 1))], 1)
    return coor, tri


def crack_family(n: int = 16, lo: int = 1000, hi: int = 40000, seed: int = 7):
    """Sides of n square cracks (rect_grid(side) each) whose element counts
    are drawn log-uniformly in [lo, hi] with a seeded generator; each side is
    the rounded square root of its count, at least 4."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    return [max(4, int(round(np.sqrt(c)))) for c in counts]
