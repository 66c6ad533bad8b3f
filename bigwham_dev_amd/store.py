"""On-disk H-matrix: save a built operator once, replay it without the host build.

The reference writes an HDF5 file with the groups of ``Hmat::writeToFile``
(src/hmat/hmatrix/hmat.cc:749-840): attributes ``/hmat/{dof_dimension,m,n}``,
``/hmat/pattern/{n_FRB,n_LRB,FRB,LRB}``, ``/hmat/permutations/{rows,cols}``,
one dataset per full-rank block ``/hmat/FRB/frb_<k>`` and one group per
low-rank block ``/hmat/LRB/lrb_<k>/{A,B}``; ``readFromFile`` (hmat.cc:842-935)
reads them back.  HDF5 is not available here, so the same content is stored
as a directory of ``.npy`` arrays that numpy memory-maps on load (no pickle):

    meta.json                  format, version, dof_dimension, m, n, n_FRB, n_LRB
    pattern_FRB.npy            (n_FRB, 6) int64: column k of the 6 x n_FRB table
    pattern_LRB.npy            (n_LRB, 6) int64 (row 5 = rank after build)
    perm_rows.npy, perm_cols.npy   permutation_0_ / permutation_1_ (int64)
    frb_data.npy, frb_offsets.npy  every FR block, column-major, concatenated
    lrb_A.npy, lrb_A_offsets.npy   every A factor ((dof m) x r, column-major)
    lrb_B.npy, lrb_B_offsets.npy   every B factor (r x (dof n), column-major)

Concatenating the blocks keeps the file count at ten for 10^5 blocks; the
offsets give each block's start (n + 1 entries).  ``FlatFile`` rebuilds the
``bw_flat_hmat`` view with pointers into the mapped arrays, so ``HipHmat``
uploads straight from the page cache.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path

import numpy as np

from . import _native as N

FORMAT = "bigwham_dev_amd flat H-matrix"
VERSION = 1


def _blocks_from_flat(f, which):
    """Column-major payloads of the blocks of a bw_flat_hmat, as numpy views."""
    d = f.dof
    out = []
    if which == "fr":
        pat = np.ctypeslib.as_array(f.fr_pattern, shape=(f.n_fr, 6)) if f.n_fr else np.zeros((0, 6), np.int64)
        for k in range(f.n_fr):
            m = d * (pat[k, 3] - pat[k, 1])
            n = d * (pat[k, 4] - pat[k, 2])
            out.append(np.ctypeslib.as_array(f.fr_blocks[k], shape=(m * n,)))
        return pat, out
    pat = np.ctypeslib.as_array(f.lr_pattern, shape=(f.n_lr, 6)) if f.n_lr else np.zeros((0, 6), np.int64)
    A, B = [], []
    for k in range(f.n_lr):
        m = d * (pat[k, 3] - pat[k, 1])
        n = d * (pat[k, 4] - pat[k, 2])
        r = int(pat[k, 5])
        A.append(np.ctypeslib.as_array(f.lr_A[k], shape=(m * r,)) if r else np.zeros(0))
        B.append(np.ctypeslib.as_array(f.lr_B[k], shape=(r * n,)) if r else np.zeros(0))
    return pat, (A, B)


def _concat(parts, path: Path):
    """Write the concatenation of `parts` to a .npy without holding it twice."""
    offs = np.zeros(len(parts) + 1, dtype=np.int64)
    np.cumsum([p.size for p in parts], out=offs[1:])
    if offs[-1] == 0:
        np.save(path, np.zeros(0))
        return offs
    arr = np.lib.format.open_memmap(path, mode="w+", dtype=np.float64, shape=(int(offs[-1]),))
    for p, o in zip(parts, offs[:-1]):
        arr[o:o + p.size] = p
    arr.flush()
    del arr
    return offs


def save(hmat, path) -> Path:
    """Write a built operator (anything with ``.flat()``: HostHmat, FlatFile)."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    f = hmat.flat()
    if f.n_fr and not f.fr_blocks:
        raise ValueError("the H-matrix has no blocks yet (pattern only): build it first")
    fpat, fr = _blocks_from_flat(f, "fr")
    lpat, (A, B) = _blocks_from_flat(f, "lr")
    np.save(path / "pattern_FRB.npy", np.ascontiguousarray(fpat))
    np.save(path / "pattern_LRB.npy", np.ascontiguousarray(lpat))
    np.save(path / "perm_rows.npy", np.ctypeslib.as_array(f.perm0, shape=(f.n_row_pts,)).copy())
    np.save(path / "perm_cols.npy", np.ctypeslib.as_array(f.perm1, shape=(f.n_col_pts,)).copy())
    np.save(path / "frb_offsets.npy", _concat(fr, path / "frb_data.npy"))
    np.save(path / "lrb_A_offsets.npy", _concat(A, path / "lrb_A.npy"))
    np.save(path / "lrb_B_offsets.npy", _concat(B, path / "lrb_B.npy"))
    meta = {"format": FORMAT, "version": VERSION, "dof_dimension": int(f.dof),
            "m": int(f.dof * f.n_row_pts), "n": int(f.dof * f.n_col_pts),
            "n_FRB": int(f.n_fr), "n_LRB": int(f.n_lr)}
    tmp = path / "meta.json.tmp"
    tmp.write_text(json.dumps(meta, indent=1))
    os.replace(tmp, path / "meta.json")  # meta last: a partial write is not a file
    return path


class FlatFile:
    """A saved operator, memory-mapped; ``flat()`` gives the bw_flat_hmat that
    HipHmat / HipHmat.many accept in place of a HostHmat."""

    def __init__(self, path):
        path = Path(path)
        try:
            meta = json.loads((path / "meta.json").read_text())
        except (OSError, ValueError) as e:
            raise ValueError(f"{path}: not a saved H-matrix ({e})") from None
        if meta.get("format") != FORMAT or meta.get("version") != VERSION:
            raise ValueError(f"{path}: unsupported format {meta.get('format')!r} "
                             f"version {meta.get('version')!r}")
        def ld(name):
            try:
                return np.load(path / name, mmap_mode="r", allow_pickle=False)
            except ValueError:  # an empty array cannot be memory-mapped
                return np.load(path / name, allow_pickle=False)
        def ld_int(name):
            a = ld(name)
            if a.dtype.kind not in "iu":
                raise ValueError(f"{path}: {name} holds {a.dtype}, not integers")
            return np.ascontiguousarray(a, dtype=np.int64)

        self.meta = meta
        self.dof = d = int(meta["dof_dimension"])
        self.fr_pattern = ld_int("pattern_FRB.npy")
        self.lr_pattern = ld_int("pattern_LRB.npy")
        self.perm0 = ld_int("perm_rows.npy")
        self.perm1 = ld_int("perm_cols.npy")
        nfr, nlr = int(meta["n_FRB"]), int(meta["n_LRB"])
        if (self.fr_pattern.shape != (nfr, 6) or self.lr_pattern.shape != (nlr, 6)
                or d * self.perm0.size != meta["m"] or d * self.perm1.size != meta["n"]):
            raise ValueError(f"{path}: pattern or permutation sizes disagree with meta.json")
        def ld_payload(name):
            # the native code reads 8 bytes per entry through raw pointers:
            # only 1-D, C-contiguous little-endian float64 is accepted
            a = ld(name)
            if a.ndim != 1 or a.dtype != np.dtype("<f8") or not a.flags.c_contiguous:
                raise ValueError(f"{path}: {name} must be a 1-D little-endian float64 array "
                                 f"(found {a.dtype}, shape {a.shape})")
            return a

        self._fr = ld_payload("frb_data.npy")
        self._a = ld_payload("lrb_A.npy")
        self._b = ld_payload("lrb_B.npy")
        fo, ao, bo = (ld_int("frb_offsets.npy"), ld_int("lrb_A_offsets.npy"),
                      ld_int("lrb_B_offsets.npy"))
        fp, lp = self.fr_pattern, self.lr_pattern
        fr_sz = d * (fp[:, 3] - fp[:, 1]) * d * (fp[:, 4] - fp[:, 2])
        a_sz = d * (lp[:, 3] - lp[:, 1]) * lp[:, 5]
        b_sz = lp[:, 5] * d * (lp[:, 4] - lp[:, 2])
        for name, offs, sz, data in (("FRB", fo, fr_sz, self._fr), ("LRB A", ao, a_sz, self._a),
                                     ("LRB B", bo, b_sz, self._b)):
            if (offs.size != sz.size + 1 or offs[0] != 0 or not np.array_equal(np.diff(offs), sz)
                    or offs[-1] != data.size):
                raise ValueError(f"{path}: {name} payload sizes disagree with the pattern")
        self.size = int(meta["m"])
        self._keep = []
        self._flat = self._make_flat(fo, ao, bo)

    def _ptrs(self, data, offs, count):
        base = data.ctypes.data if data.size else 0
        arr = (C.POINTER(C.c_double) * max(count, 1))()
        for k in range(count):
            arr[k] = C.cast(base + 8 * int(offs[k]), C.POINTER(C.c_double))
        self._keep.append(arr)
        return arr

    def _make_flat(self, fo, ao, bo):
        f = N.FlatHmat()
        f.dof = self.dof
        f.n_row_pts = self.perm0.size
        f.n_col_pts = self.perm1.size
        f.perm0 = N.iptr(self.perm0)
        f.perm1 = N.iptr(self.perm1)
        f.n_fr = self.fr_pattern.shape[0]
        f.fr_pattern = N.iptr(self.fr_pattern)
        f.fr_blocks = self._ptrs(self._fr, fo, f.n_fr)
        f.n_lr = self.lr_pattern.shape[0]
        f.lr_pattern = N.iptr(self.lr_pattern)
        f.lr_A = self._ptrs(self._a, ao, f.n_lr)
        f.lr_B = self._ptrs(self._b, bo, f.n_lr)
        return f

    def flat(self):
        return self._flat

    def nb_entries(self) -> int:
        return int(self._fr.size + self._a.size + self._b.size)


def load(path) -> FlatFile:
    return FlatFile(path)
