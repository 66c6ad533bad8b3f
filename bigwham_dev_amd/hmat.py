"""Python handles over the two native libraries.

HostHmat  — host-side construction (libbwhost.so): mesh, kernel, cluster tree,
            pattern, full-rank blocks and ACA low-rank blocks; mirrors
            BigWhamIO's constructors (square and rectangular) / BuildPattern /
            BuildHierarchicalMatrix (src/io/bigwham_io.cc:41-290, :292-484,
            :487-512, :635-704).
HipHmat   — the MI355X engine (libbwhip.so): device-resident layout and the
            matvec / adjoint / batched entry points (drop-in for
            Hmat::matvec / matvecOriginal, src/hmat/hmatrix/hmat.cc:472-715,
            and HmatCuda::matvec, hmat_cuda.cc:1771-2460).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class HostHmat:
    """Host-side H-matrix (owned by libbwhost)."""

    def __init__(self, handle: int, keepalive=None):
        self._h = C.c_void_p(handle)
        self._keep = keepalive
        self._flat = None

    # ---- construction -------------------------------------------------
    @classmethod
    def from_mesh(cls, kernel: str, coor, conn, properties, n_threads: int = -1,
                  homogeneous_size_pattern: bool = False, fixed_rank: int = -1) -> "HostHmat":
        coor = np.ascontiguousarray(coor, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int64)
        props = np.ascontiguousarray(np.asarray(properties, dtype=np.float64).ravel())
        out = C.c_void_p()
        N.check_host(N.host().bwh_create(kernel.encode(), N.dptr(coor), coor.shape[0],
                                         N.iptr(conn), conn.shape[0], N.dptr(props),
                                         props.size, int(n_threads),
                                         int(bool(homogeneous_size_pattern)), int(fixed_rank),
                                         C.byref(out)))
        return cls(out.value)

    @classmethod
    def from_meshes(cls, kernel: str, coor_src, conn_src, coor_rec, conn_rec, properties,
                    n_threads: int = -1, homogeneous_size_pattern: bool = False,
                    fixed_rank: int = -1) -> "HostHmat":
        """Rectangular operator from the source mesh (columns) to the receiver
        mesh's collocation points (rows); kernel "2DS0-2DS0-H", "3DT0-3DT0-H",
        "3DR0-3DR0-H" or "3DR0-3DR0-H-mode1" (BigWhamIO rectangular
        constructor, src/io/bigwham_io.cc:292-484)."""
        cs = np.ascontiguousarray(coor_src, dtype=np.float64)
        ks = np.ascontiguousarray(conn_src, dtype=np.int64)
        cr = np.ascontiguousarray(coor_rec, dtype=np.float64)
        kr = np.ascontiguousarray(conn_rec, dtype=np.int64)
        props = np.ascontiguousarray(np.asarray(properties, dtype=np.float64).ravel())
        out = C.c_void_p()
        N.check_host(N.host().bwh_create_rect(kernel.encode(), N.dptr(cs), cs.shape[0],
                                              N.iptr(ks), ks.shape[0], N.dptr(cr), cr.shape[0],
                                              N.iptr(kr), kr.shape[0], N.dptr(props),
                                              props.size, int(n_threads),
                                              int(bool(homogeneous_size_pattern)),
                                              int(fixed_rank), C.byref(out)))
        return cls(out.value)

    @classmethod
    def gaussian(cls, n: int, alpha: float, n_threads: int = -1,
                 homogeneous_size_pattern: bool = True) -> "HostHmat":
        out = C.c_void_p()
        N.check_host(N.host().bwh_create_gaussian(int(n), float(alpha), int(n_threads),
                                                  int(bool(homogeneous_size_pattern)),
                                                  C.byref(out)))
        return cls(out.value)

    def build_pattern(self, max_leaf_size: int, eta: float) -> None:
        N.check_host(N.host().bwh_build_pattern(self._h, int(max_leaf_size), float(eta)))
        self._flat = None

    def build(self, max_leaf_size: int, eta: float, eps_aca: float) -> None:
        N.check_host(N.host().bwh_build_hmat(self._h, int(max_leaf_size), float(eta),
                                             float(eps_aca)))
        self._flat = None

    def __del__(self):
        try:
            if self._h:
                N.host().bwh_free(self._h)
                self._h = None
        except Exception:
            pass

    # ---- views ----------------------------------------------------------
    def flat(self) -> N.FlatHmat:
        if self._flat is None:
            f = N.FlatHmat()
            N.check_host(N.host().bwh_flat(self._h, C.byref(f)))
            self._flat = f
        return self._flat

    def info(self):
        dof, dim, n = C.c_int(), C.c_int(), C.c_int64()
        N.check_host(N.host().bwh_info(self._h, C.byref(dof), C.byref(dim), C.byref(n)))
        return dof.value, dim.value, n.value

    @property
    def dof(self) -> int:
        return self.info()[0]

    @property
    def size(self) -> int:
        """dof x the mesh's collocation points: the order of a square
        operator (the column count of a rectangular one; see `shape`)."""
        dof, _, n = self.info()
        return dof * n

    def shape_points(self):
        """(receiver points, source points): equal for a square operator."""
        r, c = C.c_int64(), C.c_int64()
        N.check_host(N.host().bwh_shape(self._h, C.byref(r), C.byref(c)))
        return r.value, c.value

    @property
    def shape(self):
        """(rows, columns) of the operator in DoFs."""
        d = self.dof
        r, c = self.shape_points()
        return d * r, d * c

    def collocation_points(self) -> np.ndarray:
        dof, dim, n = self.info()
        out = np.empty((n, dim))
        N.check_host(N.host().bwh_collocation_points(self._h, N.dptr(out)))
        return out

    def receiver_collocation_points(self) -> np.ndarray:
        """Collocation points of the receiver mesh (the mesh's own for a square
        operator), original order."""
        _, dim, _ = self.info()
        r, _ = self.shape_points()
        out = np.empty((r, dim))
        N.check_host(N.host().bwh_receiver_collocation_points(self._h, N.dptr(out)))
        return out

    def element_normals(self) -> np.ndarray:
        """(n_elts, dim) unit normals (BigWhamIO::GetElementNormals)."""
        _, dim, n = self.info()
        out = np.empty(n * dim)
        N.check_host(N.host().bwh_element_normals(self._h, N.dptr(out)))
        return out.reshape(n, dim)

    def rotation_matrices(self) -> np.ndarray:
        """(n_elts, dim, dim) global-to-local rotations R of every element
        (BigWhamIO::GetRotationMatrix): local = R @ global."""
        _, dim, n = self.info()
        out = np.empty(n * dim * dim)
        N.check_host(N.host().bwh_rotation_matrices(self._h, N.dptr(out)))
        return out.reshape(n, dim, dim)

    def permutation(self) -> np.ndarray:
        f = self.flat()
        return np.ctypeslib.as_array(f.perm1, shape=(f.n_col_pts,)).copy()

    def permutation_receivers(self) -> np.ndarray:
        """Row (receiver) permutation, perm0; the same array as permutation()
        for a square operator."""
        f = self.flat()
        return np.ctypeslib.as_array(f.perm0, shape=(f.n_row_pts,)).copy()

    def fr_pattern(self) -> np.ndarray:
        f = self.flat()
        if f.n_fr == 0:
            return np.zeros((6, 0), dtype=np.int64)
        return np.ctypeslib.as_array(f.fr_pattern, shape=(f.n_fr, 6)).T.copy()

    def lr_pattern(self) -> np.ndarray:
        f = self.flat()
        if f.n_lr == 0:
            return np.zeros((6, 0), dtype=np.int64)
        return np.ctypeslib.as_array(f.lr_pattern, shape=(f.n_lr, 6)).T.copy()

    def fr_block(self, k: int) -> np.ndarray:
        f = self.flat()
        p = self.fr_pattern()[:, k]
        d = f.dof
        m, n = d * (p[3] - p[1]), d * (p[4] - p[2])
        return np.ctypeslib.as_array(f.fr_blocks[k], shape=(n, m)).T.copy()

    def lr_block(self, k: int):
        f = self.flat()
        p = self.lr_pattern()[:, k]
        d = f.dof
        m, n, r = d * (p[3] - p[1]), d * (p[4] - p[2]), p[5]
        if r == 0:
            return np.zeros((m, 0)), np.zeros((0, n))
        A = np.ctypeslib.as_array(f.lr_A[k], shape=(r, m)).T.copy()
        B = np.ctypeslib.as_array(f.lr_B[k], shape=(n, r)).T.copy()
        return A, B

    def nb_entries(self) -> int:
        d = self.dof
        fp, lp = self.fr_pattern(), self.lr_pattern()
        e = int(np.sum(d * (fp[3] - fp[1]) * d * (fp[4] - fp[2])))
        e += int(np.sum(d * (lp[3] - lp[1]) * lp[5] + lp[5] * d * (lp[4] - lp[2])))
        return e

    def compression_ratio(self) -> float:
        """nbOfEntries / (rows x columns) (Hmat::compressionRatio, hmat.cc:464-467)."""
        r, c = self.shape
        return self.nb_entries() / float(r * c)

    def dense_block(self, b0: int, b1: int, m: int, n: int) -> np.ndarray:
        d = self.dof
        out = np.empty((d * n, d * m))  # column-major (d*m) x (d*n) seen transposed
        N.check_host(N.host().bwh_dense_block(self._h, b0, b1, m, n, N.dptr(out)))
        return out.T

    def dense(self, original: bool = False) -> np.ndarray:
        """Exact operator (toArray2D, matrix_generator.h:35-40), permuted or original order."""
        dof = self.dof
        r, c = self.shape_points()
        M = self.dense_block(0, 0, r, c)
        if not original:
            return M
        ar = np.arange(dof)[None, :]
        ri = (dof * self.permutation_receivers()[:, None] + ar).ravel()
        ci = (dof * self.permutation()[:, None] + ar).ravel()
        out = np.empty_like(M)
        out[np.ix_(ri, ci)] = M
        return out

    def influence(self, e_src: int, e_rec: int) -> np.ndarray:
        d = self.dof
        out = np.empty(d * d)
        N.check_host(N.host().bwh_influence(self._h, e_src, e_rec, N.dptr(out)))
        return out.reshape(d, d).T

    def max_aca_error(self) -> float:
        v = C.c_double()
        N.check_host(N.host().bwh_max_aca_error(self._h, C.byref(v)))
        return v.value

    def timings(self):
        a, b = C.c_double(), C.c_double()
        N.check_host(N.host().bwh_timings(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value


class HipHmat:
    """Device-resident H-matrix on one MI355X (owned by libbwhip)."""

    def __init__(self, host: HostHmat, precisions=("fp64",), adjoint: bool = False,
                 device: int = -1, fp32_rank_tol: float = 0.0):
        """fp32_rank_tol > 0 recompresses the low-rank blocks of the FP32 copy
        to that relative accuracy (bw_hip_set_fp32_rank_tolerance); the FP64
        copy is the host's blocks as they are."""
        mask, flags = self._mask(precisions, adjoint)
        out = C.c_void_p()
        self._host = host  # host arrays may be freed after create; kept for convenience
        lib = N.hip()
        N.check_hip(lib.bw_hip_set_fp32_rank_tolerance(float(fp32_rank_tol)))
        try:
            N.check_hip(lib.bw_hip_create(C.byref(host.flat()), mask, flags, int(device),
                                          C.byref(out)))
        finally:
            lib.bw_hip_set_fp32_rank_tolerance(0.0)
        self._finish(out, precisions, adjoint, device)
        self.fp32_rank_tol = float(fp32_rank_tol)

    @classmethod
    def many(cls, hosts, precisions=("fp64",), adjoint: bool = False, device: int = -1,
             fp32_rank_tol: float = 0.0):
        """One operator for many independent meshes: diag(H_0, ..., H_{n-1}) on
        the concatenated vectors (bw_hip_create_multi).  ``row_offsets`` /
        ``col_offsets`` give each mesh's slice of y and x.  fp32_rank_tol as
        for the constructor."""
        self = cls.__new__(cls)
        mask, flags = cls._mask(precisions, adjoint)
        flats = [h.flat() for h in hosts]
        arr = (C.POINTER(N.FlatHmat) * len(flats))(*[C.pointer(f) for f in flats])
        ro = np.zeros(len(flats) + 1, np.int64)
        co = np.zeros(len(flats) + 1, np.int64)
        out = C.c_void_p()
        lib = N.hip()
        N.check_hip(lib.bw_hip_set_fp32_rank_tolerance(float(fp32_rank_tol)))
        try:
            N.check_hip(lib.bw_hip_create_multi(arr, len(flats), mask, flags, int(device),
                                                C.byref(out), N.iptr(ro), N.iptr(co)))
        finally:
            lib.bw_hip_set_fp32_rank_tolerance(0.0)
        self._host = list(hosts)
        self._finish(out, precisions, adjoint, device)
        self.row_offsets, self.col_offsets = ro, co
        self.fp32_rank_tol = float(fp32_rank_tol)
        return self

    @classmethod
    def selection(cls, host, rows, cols, original: bool = True, precisions=("fp64",),
                  adjoint: bool = False, device: int = -1):
        """The sub-matrix H[rows, cols] of point indices (original numbering,
        or permuted with original=False) as its own device operator
        (bw_hip_create_selection); its matvec takes the same `original`."""
        self = cls.__new__(cls)
        mask, flags = cls._mask(precisions, adjoint)
        r = np.ascontiguousarray(np.asarray(rows).ravel(), dtype=np.int64)
        c = np.ascontiguousarray(np.asarray(cols).ravel(), dtype=np.int64)
        out = C.c_void_p()
        N.check_hip(N.hip().bw_hip_create_selection(
            C.byref(host.flat()), N.iptr(r), r.size, N.iptr(c), c.size,
            N.BW_ORDER_ORIGINAL if original else N.BW_ORDER_PERMUTED, mask, flags,
            int(device), C.byref(out)))
        self._host = host
        self._finish(out, precisions, adjoint, device)
        self.selection_original = bool(original)
        return self

    @staticmethod
    def _mask(precisions, adjoint):
        mask = 0
        for p in precisions:
            mask |= {"fp64": N.BW_PREC_FP64, "fp32": N.BW_PREC_FP32}[p]
        return mask, (N.BW_FLAG_ADJOINT if adjoint else 0)

    def _finish(self, out, precisions, adjoint, device):
        self._h = out
        self.precisions = tuple(precisions)
        self.adjoint = adjoint
        r, c, e, db, s64, s32 = (C.c_int64() for _ in range(6))
        N.check_hip(N.hip().bw_hip_sizes(self._h, C.byref(r), C.byref(c), C.byref(e),
                                         C.byref(db), C.byref(s64), C.byref(s32)))
        self.n_rows, self.n_cols = r.value, c.value
        self.nb_entries = e.value
        # per precision: (FR entries, A entries, B entries, total rank)
        self.entries = {}
        for name in precisions:
            v = [C.c_int64() for _ in range(4)]
            N.check_hip(N.hip().bw_hip_entries(self._h, self._prec(name), *map(C.byref, v)))
            self.entries[name] = tuple(x.value for x in v)
        self.device_bytes = db.value
        self.stream_bytes = {"fp64": s64.value, "fp32": s32.value}
        self.device = int(device)
        self.torch_device = f"cuda:{device}" if device >= 0 else "cuda"

    def close(self) -> None:
        """Free the device copy now (also done when the object is collected)."""
        if getattr(self, "_h", None):
            N.hip().bw_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _prec(p: str) -> int:
        return {"fp64": N.BW_PREC_FP64, "fp32": N.BW_PREC_FP32}[p]

    def matvec_ptr(self, x_ptr: int, y_ptr: int, original: bool = True, precision="fp64",
                   stream: int = 0, adjoint: bool = False) -> None:
        fn = N.hip().bw_hip_rmatvec if adjoint else N.hip().bw_hip_matvec
        N.check_hip(fn(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                       N.BW_ORDER_ORIGINAL if original else N.BW_ORDER_PERMUTED,
                       self._prec(precision), 1, C.c_void_p(stream)))

    def matvec(self, x, original: bool = True, precision: str = "fp64", out=None,
               adjoint: bool = False):
        """y = H x (or H^T x).  numpy in -> numpy out (host staging, like
        HmatCuda::matvec(ArrayView)); torch CUDA tensor in -> tensor out."""
        n_in, n_out = (self.n_rows, self.n_cols) if adjoint else (self.n_cols, self.n_rows)
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, dtype=np.float64)
            if x.size != n_in:
                raise ValueError(f"matvec: x has {x.size} entries, the operator takes {n_in}")
            y = np.empty(n_out) if out is None else out
            if (not isinstance(y, np.ndarray) or y.dtype != np.float64 or y.size != n_out
                    or not y.flags.c_contiguous):
                raise ValueError(f"matvec: out must be a contiguous float64 array of {n_out}")
            fn = N.hip().bw_hip_rmatvec if adjoint else N.hip().bw_hip_matvec
            N.check_hip(fn(self._h, x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p),
                           N.BW_ORDER_ORIGINAL if original else N.BW_ORDER_PERMUTED,
                           self._prec(precision), 0, None))
            return y
        import torch
        if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64):
            raise TypeError("x must be a numpy array or a float64 CUDA tensor")
        x = x.contiguous()
        if x.numel() != n_in:
            raise ValueError(f"matvec: x has {x.numel()} entries, the operator takes {n_in}")
        y = torch.empty(n_out, dtype=torch.float64, device=x.device) if out is None else out
        if (not isinstance(y, torch.Tensor) or y.dtype != torch.float64 or y.device != x.device
                or y.numel() != n_out or not y.is_contiguous()):
            raise ValueError(f"matvec: out must be a contiguous float64 tensor of {n_out} "
                             "on the device of x")
        stream = torch.cuda.current_stream(x.device).cuda_stream
        self.matvec_ptr(x.data_ptr(), y.data_ptr(), original, precision, stream, adjoint)
        return y

    def rmatvec(self, x, original: bool = True, precision: str = "fp64", out=None):
        return self.matvec(x, original, precision, out, adjoint=True)

    def matvec_batched(self, X, original: bool = True, precision: str = "fp64", out=None,
                       adjoint: bool = False):
        """Y = H X (or H^T X) for a (n, k) float64 CUDA tensor, 32 columns per
        MFMA tile.  X may have any positive strides (row-major, column-major,
        a column slice) and is read in place; Y is row-major like torch's
        default unless ``out`` (an (n_out, k) float64 tensor) is given."""
        import torch
        n_in, n_out = (self.n_rows, self.n_cols) if adjoint else (self.n_cols, self.n_rows)
        if X.dim() != 2 or X.shape[0] != n_in or X.dtype != torch.float64 or not X.is_cuda:
            raise ValueError(f"matvec_batched: X must be a ({n_in}, k) float64 CUDA tensor")
        k = X.shape[1]
        if min(X.stride()) < 1:  # broadcast (stride 0) views have no distinct entries
            X = X.contiguous()
        Y = out if out is not None else torch.empty((n_out, k), dtype=torch.float64,
                                                     device=X.device)
        if Y.shape != (n_out, k) or Y.dtype != torch.float64 or Y.device != X.device:
            raise ValueError(f"matvec_batched: out must be a ({n_out}, {k}) float64 tensor "
                             "on the device of X")
        stream = torch.cuda.current_stream(X.device).cuda_stream
        N.check_hip(N.hip().bw_hip_matvec_batched_strided(
            self._h, int(adjoint), k, C.c_void_p(X.data_ptr()), X.stride(0), X.stride(1),
            C.c_void_p(Y.data_ptr()), Y.stride(0), Y.stride(1),
            N.BW_ORDER_ORIGINAL if original else N.BW_ORDER_PERMUTED,
            self._prec(precision), C.c_void_p(stream)))
        return Y

    def rmatvec_batched(self, X, original: bool = True, precision: str = "fp64"):
        return self.matvec_batched(X, original, precision, adjoint=True)

    def diagonal(self, original: bool = True) -> np.ndarray:
        d = np.empty(max(self.n_rows, self.n_cols))
        N.check_hip(N.hip().bw_hip_diagonal(self._h, N.dptr(d),
                                            N.BW_ORDER_ORIGINAL if original else N.BW_ORDER_PERMUTED))
        return d

    def graph(self, x, y=None, original: bool = True, precision: str = "fp64",
              adjoint: bool = False) -> "MatvecGraph":
        """Capture y = H x (or H^T x) on these device tensors as one HIP graph."""
        return MatvecGraph(self, x, y, original, precision, adjoint)

    def graph_batched(self, X, Y=None, original: bool = True, precision: str = "fp64",
                      adjoint: bool = False) -> "MatvecGraph":
        """Capture Y = H X (or H^T X) for an (n, k) float64 CUDA tensor X (any
        positive strides, as matvec_batched) as one HIP graph
        (bw_hip_graph_create_batched); Y defaults to a row-major (n_out, k)."""
        return MatvecGraph(self, X, Y, original, precision, adjoint, batched=True)

    def release_stream(self, stream: int | None = None) -> None:
        """Free the device scratch kept for a stream (bw_hip_stream_release)."""
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        N.check_hip(N.hip().bw_hip_stream_release(self._h, C.c_void_p(stream)))

    def set_unroll(self, fp64: int = 0, fp32: int = 0) -> None:
        N.check_hip(N.hip().bw_hip_set_unroll(self._h, int(fp64), int(fp32)))

    def set_unroll_pass(self, precision: str, pass_: int, unroll: int) -> None:
        """Unroll depth of one pass (1 or 2) of one precision."""
        N.check_hip(N.hip().bw_hip_set_unroll_pass(self._h, self._prec(precision), int(pass_),
                                                   int(unroll)))

    def set_mfma_schedule(self, schedule: int, precision: str = "fp32") -> None:
        """Schedule (0-15) of the batched MFMA kernel of one precision."""
        N.check_hip(N.hip().bw_hip_set_mfma_schedule(self._h, self._prec(precision),
                                                     int(schedule)))

    def set_timing(self, enable: bool = True) -> None:
        N.check_hip(N.hip().bw_hip_set_timing(self._h, int(enable)))

    def timing_stats(self):
        """Average ms per matvec since set_timing(): dict of the five kernels
        (gather, p1_slab, p1_reduce, p2_slab, p2_reduce), 'total' and 'count'."""
        avg = np.zeros(6)
        n = C.c_int64()
        N.check_hip(N.hip().bw_hip_timing_stats(self._h, N.dptr(avg), C.byref(n)))
        keys = ["gather", "p1_slab", "p1_reduce", "p2_slab", "p2_reduce", "total"]
        out = dict(zip(keys, avg.tolist()))
        out["count"] = n.value
        return out


class MatvecGraph:
    """A matvec (or a batched product) on fixed device operands, replayed with
    one graph launch (bw_hip_graph_create[_batched] / bw_hip_graph_launch).
    The graph owns its scratch; two replays of one graph must be ordered."""

    def __init__(self, dev: HipHmat, x, y, original, precision, adjoint, batched=False):
        import torch
        n_in, n_out = (dev.n_rows, dev.n_cols) if adjoint else (dev.n_cols, dev.n_rows)
        order = N.BW_ORDER_ORIGINAL if original else N.BW_ORDER_PERMUTED
        out = C.c_void_p()
        if batched:
            X = x
            if (not isinstance(X, torch.Tensor) or X.dim() != 2 or X.shape[0] != n_in
                    or X.dtype != torch.float64 or not X.is_cuda or min(X.stride()) < 1):
                raise ValueError(f"graph_batched: X must be a ({n_in}, k) float64 CUDA tensor "
                                 "with positive strides")
            k = X.shape[1]
            if y is None:
                y = torch.empty((n_out, k), dtype=torch.float64, device=X.device)
            if (not isinstance(y, torch.Tensor) or y.shape != (n_out, k)
                    or y.dtype != torch.float64 or y.device != X.device):
                raise ValueError(f"graph_batched: Y must be a ({n_out}, {k}) float64 tensor "
                                 "on the device of X")
            N.check_hip(N.hip().bw_hip_graph_create_batched(
                dev._h, int(adjoint), k, C.c_void_p(X.data_ptr()), X.stride(0), X.stride(1),
                C.c_void_p(y.data_ptr()), y.stride(0), y.stride(1), order,
                HipHmat._prec(precision), C.byref(out)))
        else:
            if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64
                    and x.is_contiguous()):
                raise TypeError("x must be a contiguous float64 CUDA tensor")
            if x.numel() != n_in:
                raise ValueError(f"graph: x has {x.numel()} entries, the operator takes {n_in}")
            if y is None:
                y = torch.empty(n_out, dtype=torch.float64, device=x.device)
            if (not isinstance(y, torch.Tensor) or y.dtype != torch.float64
                    or y.device != x.device or y.numel() != n_out or not y.is_contiguous()):
                raise ValueError(f"graph: y must be a contiguous float64 tensor of {n_out} "
                                 "on the device of x")
            N.check_hip(N.hip().bw_hip_graph_create(
                dev._h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), order,
                HipHmat._prec(precision), int(adjoint), C.byref(out)))
        self.dev, self.x, self.y = dev, x, y
        self._g = out

    def launch(self, stream: int | None = None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(self.x.device).cuda_stream
        N.check_hip(N.hip().bw_hip_graph_launch(self._g, C.c_void_p(stream)))
        return self.y

    def close(self) -> None:
        if getattr(self, "_g", None):
            N.hip().bw_hip_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def full_blocks(hmat, original: bool = True, rows=None, cols=None):
    """Entries of the full-rank blocks as (values, row, col) DoF coordinate
    arrays (bw_flat_full_blocks): the whole matrix, or the entries inside a
    point selection indexed in that selection.  `hmat` is anything with
    .flat() (HostHmat, store.FlatFile)."""
    f = hmat.flat()
    order = N.BW_ORDER_ORIGINAL if original else N.BW_ORDER_PERMUTED
    if (rows is None) != (cols is None):
        raise ValueError("give both rows and cols, or neither")
    if rows is not None:
        r = np.ascontiguousarray(np.asarray(rows).ravel(), dtype=np.int64)
        c = np.ascontiguousarray(np.asarray(cols).ravel(), dtype=np.int64)
        sel = (N.iptr(r), r.size, N.iptr(c), c.size)
    else:
        sel = (None, 0, None, 0)
    lib = N.hip()
    n = C.c_int64()

    def check(rc):
        if rc != N.BW_OK:
            raise N.BigWhamError(rc, lib.bw_flat_full_blocks_last_error().decode())
    check(lib.bw_flat_full_blocks(C.byref(f), order, *sel, C.byref(n), None, None, None))
    vals = np.empty(n.value)
    ri = np.empty(n.value, dtype=np.int64)
    ci = np.empty(n.value, dtype=np.int64)
    check(lib.bw_flat_full_blocks(C.byref(f), order, *sel, C.byref(n), N.dptr(vals),
                                  N.iptr(ri), N.iptr(ci)))
    return vals, ri, ci


def set_stacking(stack_rows: int = 64) -> None:
    """Pass-1 stacking threshold for the next HipHmat (bw_hip_set_stacking)."""
    N.check_hip(N.hip().bw_hip_set_stacking(int(stack_rows)))


def set_tuning(wi_bytes_fp64: int = 0, wi_bytes_fp32: int = 0) -> None:
    N.check_hip(N.hip().bw_hip_set_tuning(int(wi_bytes_fp64), int(wi_bytes_fp32)))
