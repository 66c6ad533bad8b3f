"""Random H-matrix data in the reference's flat layout (test input).

A rectangular operator (receiver and source point sets of different sizes,
as the reference's Hmat between two meshes): the row and column points are
cut into clusters of random size, and every cluster pair is a full-rank
block or a low-rank A·B block of random rank, 0 to 70 (empty blocks and
ranks over 64 included).  Pattern columns are (spot, i0, j0, iend, jend,
flag) in point units with iend / jend exclusive and flag = rank for low-rank
blocks (h_pattern.h:21-33, hmat.cc:390); blocks are column-major with leading
dimension = rows (Array2D.h:840).  `dense()` assembles the operator in numpy
from the same blocks, independently of the oracle.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from bigwham_dev_amd import _native as N


class SyntheticHmat:
    def __init__(self, n_row_pts: int, n_col_pts: int, dof: int, seed: int,
                 p_full: float = 0.3, max_rank: int = 70):
        rng = np.random.default_rng(seed)

        def clusters(n):
            cuts = [0]
            while cuts[-1] < n:
                cuts.append(min(n, cuts[-1] + int(rng.integers(3, 40))))
            return list(zip(cuts[:-1], cuts[1:]))

        self.dof = dof
        self.n_row_pts, self.n_col_pts = n_row_pts, n_col_pts
        fr_pat, self.fr, lr_pat, self.A, self.B = [], [], [], [], []
        spot = 0
        for i0, i1 in clusters(n_row_pts):
            for j0, j1 in clusters(n_col_pts):
                m, n = dof * (i1 - i0), dof * (j1 - j0)
                if rng.random() < p_full:
                    fr_pat.append((spot, i0, j0, i1, j1, 0))
                    self.fr.append(np.asfortranarray(rng.uniform(-1, 1, (m, n))))
                else:
                    r = int(rng.integers(0, max_rank + 1))
                    lr_pat.append((spot, i0, j0, i1, j1, r))
                    self.A.append(np.asfortranarray(rng.uniform(-1, 1, (m, r))))
                    self.B.append(np.asfortranarray(rng.uniform(-1, 1, (r, n))))
                spot += 1
        # 6 x n column-major = n x 6 row-major
        self.fr_pattern = np.ascontiguousarray(np.array(fr_pat, dtype=np.int64).reshape(-1, 6))
        self.lr_pattern = np.ascontiguousarray(np.array(lr_pat, dtype=np.int64).reshape(-1, 6))
        self.perm0 = rng.permutation(n_row_pts).astype(np.int64)
        self.perm1 = rng.permutation(n_col_pts).astype(np.int64)
        self.size = dof * n_row_pts  # rows, as HostHmat.size for square operators
        self._flat = None

    def flat(self) -> N.FlatHmat:
        if self._flat is None:
            P = C.POINTER(C.c_double)
            ptr = lambda a: a.ctypes.data_as(P) if a.size else P()
            self._fr_ptrs = (P * max(1, len(self.fr)))(*[ptr(a) for a in self.fr])
            self._a_ptrs = (P * max(1, len(self.A)))(*[ptr(a) for a in self.A])
            self._b_ptrs = (P * max(1, len(self.B)))(*[ptr(a) for a in self.B])
            f = N.FlatHmat()
            f.dof = self.dof
            f.n_row_pts, f.n_col_pts = self.n_row_pts, self.n_col_pts
            f.perm0 = N.iptr(self.perm0)
            f.perm1 = N.iptr(self.perm1)
            f.n_fr = len(self.fr)
            f.fr_pattern = N.iptr(self.fr_pattern)
            f.fr_blocks = C.cast(self._fr_ptrs, C.POINTER(P))
            f.n_lr = len(self.A)
            f.lr_pattern = N.iptr(self.lr_pattern)
            f.lr_A = C.cast(self._a_ptrs, C.POINTER(P))
            f.lr_B = C.cast(self._b_ptrs, C.POINTER(P))
            self._flat = f
        return self._flat

    def dense(self, original: bool = False) -> np.ndarray:
        d = self.dof
        H = np.zeros((d * self.n_row_pts, d * self.n_col_pts))
        for (_, i0, j0, i1, j1, _), F in zip(self.fr_pattern, self.fr):
            H[d * i0:d * i1, d * j0:d * j1] += F
        for (_, i0, j0, i1, j1, _), A, B in zip(self.lr_pattern, self.A, self.B):
            H[d * i0:d * i1, d * j0:d * j1] += A @ B
        if not original:
            return H
        # y_orig[d perm0[i] + a] = y[d i + a], x_perm[d j + b] = x_orig[d perm1[j] + b]
        ri = (d * self.perm0[:, None] + np.arange(d)).ravel()
        ci = (d * self.perm1[:, None] + np.arange(d)).ravel()
        Ho = np.empty_like(H)
        Ho[np.ix_(ri, ci)] = H
        return Ho
