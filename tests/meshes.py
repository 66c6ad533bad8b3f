"""Synthetic meshes of the benchmark configurations (seeded, deterministic)."""
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
