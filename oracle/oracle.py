"""ctypes wrapper of the CPU oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the expected result and as the host-core baseline;
never by the bigwham_dev_amd package.  Every function states the reference
routine it restates (see hmat_oracle.c).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None


def _bind(path) -> C.CDLL:
    _lib = C.CDLL(str(path))
    vp = C.c_void_p
    d = C.POINTER(C.c_double)
    _lib.oracle_matvec.argtypes = [vp, d, d, C.c_int]
    _lib.oracle_matvec_original.argtypes = [vp, d, d, C.c_int]
    _lib.oracle_rmatvec.argtypes = [vp, d, d]
    _lib.oracle_diagonal.argtypes = [vp, d]
    _lib.oracle_diagonal_original.argtypes = [vp, d]
    _lib.oracle_dense.argtypes = [vp, d]
    _lib.oracle_nb_entries.argtypes = [vp]
    _lib.oracle_nb_entries.restype = C.c_int64
    _lib.oracle_exchange_sort.argtypes = [d, C.c_int, C.POINTER(C.c_int64), C.c_int64,
                                          C.c_int64, C.c_int]
    _lib.oracle_exchange_sort.restype = None
    return _lib


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = HERE / "lib" / "liboracle.so"
        if not path.exists():
            import sys
            sys.path.insert(0, str(HERE))
            import build_oracle
            build_oracle.build_oracle()
        _lib = _bind(path)
    return _lib


def native_lib(out_dir) -> C.CDLL:
    """The oracle compiled for this host's CPU (-O3 -march=native), loaded
    from out_dir: the host-core baseline of bench.py."""
    import sys
    sys.path.insert(0, str(HERE))
    import build_oracle
    return _bind(build_oracle.build_oracle_native(out_dir))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _sizes(flat):
    d = flat.dof
    return d * flat.n_row_pts, d * flat.n_col_pts


def matvec(flat, x: np.ndarray, n_threads: int = 0) -> np.ndarray:
    """Hmat<T>::matvec (hmat.cc:472-537), permuted order."""
    nr, nc = _sizes(flat)
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert x.shape == (nc,)
    y = np.empty(nr)
    lib().oracle_matvec(C.byref(flat), _dp(x), _dp(y), int(n_threads))
    return y


def matvec_original(flat, x: np.ndarray, n_threads: int = 0, use=None) -> np.ndarray:
    """Hmat<T>::matvecOriginal (hmat.cc:572-635).  use: another build of the
    oracle (native_lib) instead of the default one."""
    nr, nc = _sizes(flat)
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert x.shape == (nc,)
    y = np.empty(nr)
    (use or lib()).oracle_matvec_original(C.byref(flat), _dp(x), _dp(y), int(n_threads))
    return y


def rmatvec(flat, x: np.ndarray) -> np.ndarray:
    """H^T x in permuted order (same blocks, transposed)."""
    nr, nc = _sizes(flat)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(nc)
    lib().oracle_rmatvec(C.byref(flat), _dp(x), _dp(y))
    return y


def rmatvec_original(flat, x: np.ndarray) -> np.ndarray:
    """P1^T H^T P0 x: gather by perm0, adjoint, scatter by perm1."""
    d = flat.dof
    p0 = np.ctypeslib.as_array(flat.perm0, shape=(flat.n_row_pts,))
    p1 = np.ctypeslib.as_array(flat.perm1, shape=(flat.n_col_pts,))
    i0 = (d * p0[:, None] + np.arange(d)).ravel()
    i1 = (d * p1[:, None] + np.arange(d)).ravel()
    z = np.asarray(x, dtype=np.float64)[i0]
    y = rmatvec(flat, z)
    out = np.empty_like(y)
    out[i1] = y
    return out


def _inverse(perm):
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    return inv


def selection_matvec(flat, rows, cols, x, original: bool = True, adjoint: bool = False):
    """HmatSelection::matvecOriginal / matvec (hmat_selection.cc:188-257):
    scatter x into a zero permuted vector at the selected columns (a repeated
    column keeps its last value), multiply by the whole H, gather the
    selected rows.  Point indices are original (through the inverse
    permutations) or permuted.  adjoint: the same for H^T, rows and columns
    swapping roles."""
    d = flat.dof
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    p0 = np.ctypeslib.as_array(flat.perm0, shape=(flat.n_row_pts,))
    p1 = np.ctypeslib.as_array(flat.perm1, shape=(flat.n_col_pts,))
    pr = _inverse(p0)[rows] if original else rows
    pc = _inverse(p1)[cols] if original else cols
    src, dst = (pr, pc) if adjoint else (pc, pr)
    x = np.asarray(x, dtype=np.float64).reshape(-1, d)
    n_in = d * (flat.n_row_pts if adjoint else flat.n_col_pts)
    xf = np.zeros((n_in // d, d))
    for i, p in enumerate(src):  # in order: the last occurrence wins
        xf[p] = x[i]
    yf = (rmatvec if adjoint else matvec)(flat, xf.ravel()).reshape(-1, d)
    return yf[dst].ravel()


def diagonal(flat, original: bool = False) -> np.ndarray:
    """Hmat<T>::diagonal / diagonalOriginal (hmat.cc:237-285)."""
    d = flat.dof
    n = d * max(flat.n_row_pts, flat.n_col_pts)
    out = np.empty(n)
    if original:
        lib().oracle_diagonal_original(C.byref(flat), _dp(out))
    else:
        lib().oracle_diagonal(C.byref(flat), _dp(out))
    return out


def dense(flat) -> np.ndarray:
    """H assembled from its blocks (FR and A*B), permuted order."""
    nr, nc = _sizes(flat)
    out = np.empty((nc, nr))
    lib().oracle_dense(C.byref(flat), _dp(out))
    return out.T.copy()


def nb_entries(flat) -> int:
    """Hmat<T>::nbOfEntries (hmat.cc:299-312)."""
    return int(lib().oracle_nb_entries(C.byref(flat)))


def exchange_sort(node: np.ndarray, perm: np.ndarray, b: int, e: int, dmax: int):
    """The literal O(n^2) exchange sort of cluster.cc:516-533 (in place)."""
    assert node.flags.c_contiguous and node.dtype == np.float64
    assert perm.dtype == np.int64
    lib().oracle_exchange_sort(_dp(node), node.shape[1],
                               perm.ctypes.data_as(C.POINTER(C.c_int64)), b, e, dmax)
