"""BEMatrix: the Python face of the reference (bigwham4py.BEMatrix,
interfaces/python/bigwham4py/bigwham4py.py:49-485) on the MI355X engine.

Same constructor arguments, same LinearOperator semantics (shape (N, N),
float64, matvec in the original DoF ordering), same helper methods for the
matvec path: getCompression, getPermutation, getCollocationPoints,
getSpatialDimension, H_diag, H_jacobi_prec, H_ILU_prec, _getFullBlocks,
get_diagonal, getMatvecTime, get_max_error_ACA, _getPattern,
write_hmatrix, plotPattern, get_element_normals, get_rotation_matrix,
convert_to_global / convert_to_local, and sub-matrix selection by indexing
(M[np.ix_(rows, cols)]).  Construction stays on the host (libbwhost);
the matvec runs on the GPU (libbwhip).  `useCuda` is accepted for signature
compatibility; there is no CPU matvec fallback in this package.

Extras of this engine: `precision="fp32"` for FP32-factor matvecs,
`matvec_torch(x)` on device tensors, `rmatvec` (H^T), and `solve()`
(FP64 or mixed-precision GMRES on the device).

BEMatrixRectangular mirrors bigwham4py.BEMatrixRectangular (bigwham4py.py:
491-708): an operator from a source mesh to the collocation points of a
receiver mesh, shape (N_receivers, N_sources).  The reference builds it for
its CPU matvec only (bigwham_io.cc:310-314 aborts with useCuda); here its
matvec, H^T and batched products run on the GPU like the square operator's.
"""
from __future__ import annotations

import os
import time

import numpy as np
from scipy.sparse import csc_matrix, diags
from scipy.sparse.linalg import LinearOperator, spilu

from .hmat import HipHmat, HostHmat, full_blocks

kernels_id = ["2DS0-H", "3DT0-H", "3DR0-H", "3DR0-H-mode1"]
kernels_id_rect = ["2DS0-2DS0-H", "3DT0-3DT0-H", "3DR0-3DR0-H", "3DR0-3DR0-H-mode1"]


class BEMatrix(LinearOperator):
    def __init__(self, kernel: str, coor: np.ndarray, conn: np.ndarray, properties: np.ndarray,
                 max_leaf_size: int = 32, eta: float = 3.0, eps_aca: float = 1.0e-3,
                 n_openMP_threads: int = 8, n_GPUs: int = 1, directly_build: bool = True,
                 verbose: bool = True, homogeneous_size_pattern: bool = False,
                 fixed_rank: int = -1, useCuda: bool = True, precisions=("fp64", "fp32"),
                 adjoint: bool = True, device: int = -1):
        if kernel not in kernels_id:
            raise ValueError(f"Invalid kernel : {kernel}, available kernels are : "
                             f"[{', '.join(kernels_id)}]")
        if n_GPUs != 1:
            raise ValueError("this engine drives one GPU per BEMatrix")
        self.kernel_ = kernel
        self.properties_ = np.asarray(properties, dtype=float)
        self.max_leaf_size_ = int(max_leaf_size)
        self.eta_ = float(eta)
        self.eps_aca_ = float(eps_aca)
        self.n_openMP_threads_ = n_openMP_threads
        self.verbose_ = verbose
        self._precisions = tuple(precisions)
        self._adjoint = adjoint
        self._device = device
        self.H_ = HostHmat.from_mesh(kernel, coor, conn, self.properties_,
                                     n_threads=n_openMP_threads,
                                     homogeneous_size_pattern=homogeneous_size_pattern,
                                     fixed_rank=fixed_rank)
        self.dev_ = None
        self.built_ = False
        self.precision = "fp64"
        if directly_build:
            self._build()
        else:
            self.H_.build_pattern(self.max_leaf_size_, self.eta_)
            super().__init__(float, (0, 0))

    def _build(self):
        if self.built_:
            return
        t0 = time.perf_counter()
        self.H_.build(self.max_leaf_size_, self.eta_, self.eps_aca_)
        self.dev_ = HipHmat(self.H_, precisions=self._precisions, adjoint=self._adjoint,
                            device=self._device)
        self.hmat_time_ = time.perf_counter() - t0
        n = self.H_.size
        self.matvec_size_ = n
        self.shape_ = (n, n)
        super().__init__(float, self.shape_)
        self.built_ = True
        if self.verbose_:
            print(f"BEMatrix {self.kernel_}: N={n}, compression ratio "
                  f"{self.getCompression():.5f}, build {self.hmat_time_:.2f} s")

    # ---- LinearOperator ---------------------------------------------------
    def _matvec(self, v: np.ndarray) -> np.ndarray:
        return self.dev_.matvec(np.asarray(v, dtype=np.float64).ravel(), original=True,
                                precision=self.precision)

    def _rmatvec(self, v: np.ndarray) -> np.ndarray:
        return self.dev_.rmatvec(np.asarray(v, dtype=np.float64).ravel(), original=True,
                                 precision=self.precision)

    def matvec_torch(self, x, precision: str | None = None):
        return self.dev_.matvec(x, original=True, precision=precision or self.precision)

    # ---- reference helpers -------------------------------------------------
    def getCompression(self) -> float:
        self._build()
        return self.H_.compression_ratio()

    def getStorageRequirement(self) -> int:
        return self.H_.nb_entries() * 8

    def getGPUStorageRequirement(self) -> int:
        return self.dev_.device_bytes

    def getPermutation(self) -> np.ndarray:
        return self.H_.permutation()

    def getCollocationPoints(self) -> np.ndarray:
        return self.H_.collocation_points()

    getMeshCollocationPoints = getCollocationPoints

    def get_omp_threads(self) -> int:
        """Host threads of the H-matrix construction (the matvec runs on the GPU)."""
        n = self.n_openMP_threads_
        return n if n and n > 0 else (os.cpu_count() or 1)

    def write_hmatrix(self, filename: str) -> int:
        """Save the built operator (Hmat::writeToFile content; as a directory
        of memory-mappable .npy arrays, see bigwham_dev_amd.store)."""
        from . import store
        self._build()
        store.save(self.H_, filename)
        return 0

    def get_element_normals(self) -> np.ndarray:
        """[n0x, n0y, (n0z), n1x, ...] as the reference returns them."""
        return self.H_.element_normals().ravel()

    def get_rotation_matrix(self) -> np.ndarray:
        """Row-major rotation matrix of every element, concatenated."""
        return self.H_.rotation_matrices().ravel()

    def convert_to_global(self, x_local: np.ndarray) -> np.ndarray:
        """Per-element vector from the element frame to the global frame,
        v = R^T x (Mesh::ConvertToGlobal, src/core/mesh.h:91-106)."""
        R = self.H_.rotation_matrices()
        x = np.asarray(x_local, dtype=np.float64).reshape(R.shape[0], R.shape[1])
        return np.einsum("eji,ej->ei", R, x).ravel()

    def convert_to_local(self, x_global: np.ndarray) -> np.ndarray:
        """Per-element vector from the global frame to the element frame,
        v = R x (Mesh::ConvertToLocal, src/core/mesh.h:108-125)."""
        R = self.H_.rotation_matrices()
        x = np.asarray(x_global, dtype=np.float64).reshape(R.shape[0], R.shape[1])
        return np.einsum("eij,ej->ei", R, x).ravel()

    def isCudaAvailable(self) -> bool:
        """True: the matvec of this engine always runs on the GPU."""
        return True

    def getSpatialDimension(self) -> int:
        return self.H_.info()[1]

    def _getPattern(self) -> np.ndarray:
        """Rows (i_begin, j_begin, i_end, j_end, FLAG, entry size) as
        BigWhamIO::GetHPattern: FLAG 0 full rank, 1 low rank."""
        fp, lp = self.H_.fr_pattern(), self.H_.lr_pattern()
        rows = [np.stack([fp[1], fp[2], fp[3], fp[4], np.zeros_like(fp[1]),
                          (fp[3] - fp[1]) * (fp[4] - fp[2])], 1),
                np.stack([lp[1], lp[2], lp[3], lp[4], np.ones_like(lp[1]), lp[5]], 1)]
        return np.concatenate(rows, 0)

    def plotPattern(self, plot_index: bool = False):
        """Figure of the block pattern (point indices, row 0 at the top):
        full-rank blocks in one colour, low-rank blocks in another, optionally
        numbered in pattern order within each kind.  Returns the figure."""
        import matplotlib.pyplot as plt
        from matplotlib.collections import PatchCollection
        from matplotlib.patches import Rectangle

        pat = self._getPattern()
        n = int(pat[:, 2].max())
        fig, ax = plt.subplots()
        rects, kinds, seen = [], [], [0, 0]
        for i0, j0, i1, j1, flag, _ in pat:
            rects.append(Rectangle((j0, n - i1), j1 - j0, i1 - i0))
            kinds.append(flag)
            if plot_index:
                ax.text((j0 + j1) / 2, n - (i0 + i1) / 2, str(seen[flag]), ha="center",
                        va="center", fontsize=7)
            seen[flag] += 1
        coll = PatchCollection(rects, cmap="coolwarm", edgecolor="black", linewidth=0.3,
                               alpha=0.5)
        coll.set_array(np.asarray(kinds, dtype=float))
        ax.add_collection(coll)
        ax.set_xlim(0, n)
        ax.set_ylim(0, n)
        ax.set_aspect("equal")
        ax.set_title(f"{seen[0]} full-rank, {seen[1]} low-rank blocks")
        return fig

    def _getFullBlocks(self, original_order: bool = True, keep_ratio: float = 1.0) -> csc_matrix:
        """Entries of the full-rank blocks as a sparse matrix
        (bigwham4py.py:252-276): all of them, or the largest keep_ratio
        fraction by magnitude."""
        if not (0 < keep_ratio <= 1):
            raise ValueError("ratio must be a float between 0 and 1.")
        sel = getattr(self, "_selection", None)
        if sel is None:
            val, row, col = full_blocks(self.H_, original=original_order)
        else:
            val, row, col = full_blocks(self.H_, original=True, rows=sel[0], cols=sel[1])
        if keep_ratio < 1.0:
            k = int(np.ceil(keep_ratio * val.size))
            if k == 0:
                return csc_matrix(self.shape_)
            thr = np.partition(np.abs(val), -k)[-k]
            m = np.abs(val) >= thr
            val, row, col = val[m], row[m], col[m]
        return csc_matrix((val, (row, col)), shape=self.shape_)

    def H_ILU_prec(self, fill_factor=1, drop_tol=1e-3, keep_ratio=1.0) -> LinearOperator:
        """ILU (scipy spilu) of the full-rank blocks (bigwham4py.py:345-361)."""
        self._build()
        ilu = spilu(self._getFullBlocks(keep_ratio=keep_ratio), fill_factor=fill_factor,
                    drop_tol=drop_tol)
        return LinearOperator(self.shape_, ilu.solve)

    def __getitem__(self, indices):
        """Sub-matrix selection (bigwham4py.py:462-483): ``M[np.ix_(rows, cols)]``
        or ``M[rows, cols]`` with point indices in the original numbering.  The
        result is a BEMatrix whose matvec runs on the GPU over the blocks that
        meet the selection only."""
        if not isinstance(indices, tuple) or len(indices) != 2:
            raise IndexError("Subsetting requires indices of the form (np.ix_(row_indices, col_indices))")
        rows = np.asarray(indices[0]).ravel().astype(np.int64)
        cols = np.asarray(indices[1]).ravel().astype(np.int64)
        new = object.__new__(BEMatrix)
        new.__dict__ = self.__dict__.copy()
        new.dev_ = HipHmat.selection(self.H_, rows, cols, original=True,
                                     precisions=self._precisions, adjoint=self._adjoint,
                                     device=self._device)
        new._selection = (rows, cols)
        new.shape_ = (new.dev_.n_rows, new.dev_.n_cols)
        new.matvec_size_ = new.shape_[0]
        LinearOperator.__init__(new, float, new.shape_)
        return new

    def get_max_error_ACA(self) -> float:
        return self.H_.max_aca_error()

    def get_diagonal(self) -> np.ndarray:
        if getattr(self, "_selection", None) is not None:
            raise NotImplementedError("the diagonal of a selection is not defined")
        return self.dev_.diagonal(original=True)

    def H_diag(self) -> np.ndarray:
        return self.get_diagonal()

    def H_jacobi_prec(self):
        return diags(1.0 / self.H_diag(), dtype=np.float64)

    def getMatvecTime(self, N_matvec: int = 100) -> float:
        x = np.ones(self.shape[0])
        self._matvec(x)
        t0 = time.time()
        for _ in range(N_matvec):
            self._matvec(x)
        return (time.time() - t0) / N_matvec

    def solve(self, b, tol: float = 1e-8, mixed: bool = True, restart: int = 30):
        """GMRES on the device (Jacobi preconditioned); returns (x, SolveInfo)."""
        import torch
        from . import solvers
        bt = b if isinstance(b, torch.Tensor) else torch.from_numpy(
            np.asarray(b, dtype=np.float64)).to("cuda")
        if mixed:
            return solvers.solve_mixed(self.dev_, bt, tol=tol, restart=restart)
        return solvers.solve_fp64(self.dev_, bt, tol=tol, restart=restart)


class BEMatrixRectangular(LinearOperator):
    """Rectangular operator between two meshes (bigwham4py.py:491-708): the
    traction at the receiver mesh's collocation points, in the receivers'
    local frames, due to displacement discontinuities on the source mesh's
    elements.  Same constructor arguments as the reference; `useCuda` is
    accepted for signature compatibility (the matvec always runs on the GPU).
    """

    def __init__(self, kernel: str, coor_src: np.ndarray, conn_src: np.ndarray,
                 coor_rec: np.ndarray, conn_rec: np.ndarray, properties: np.ndarray,
                 max_leaf_size: int = 100, eta: float = 3.0, eps_aca: float = 1.0e-3,
                 n_openMP_threads: int = 8, n_GPUs: int = 1, directly_build: bool = True,
                 verbose: bool = True, homogeneous_size_pattern: bool = False,
                 useCuda: bool = False, fixed_rank: int = -1, precisions=("fp64", "fp32"),
                 adjoint: bool = True, device: int = -1):
        if kernel not in kernels_id_rect:
            raise ValueError(f"Invalid kernel : {kernel}, available kernels are : "
                             f"[{', '.join(kernels_id_rect)}]")
        if n_GPUs != 1:
            raise ValueError("this engine drives one GPU per BEMatrixRectangular")
        self.kernel_ = kernel
        self.properties_ = np.asarray(properties, dtype=float)
        self.max_leaf_size_ = int(max_leaf_size)
        self.eta_ = float(eta)
        self.eps_aca_ = float(eps_aca)
        self.n_openMP_threads_ = n_openMP_threads
        self.verbose_ = verbose
        self._precisions = tuple(precisions)
        self._adjoint = adjoint
        self._device = device
        self.H_ = HostHmat.from_meshes(kernel, coor_src, conn_src, coor_rec, conn_rec,
                                       self.properties_, n_threads=n_openMP_threads,
                                       homogeneous_size_pattern=homogeneous_size_pattern,
                                       fixed_rank=fixed_rank)
        self.dev_ = None
        self.built_ = False
        self.precision = "fp64"
        self.shape_ = self.H_.shape
        self.matvec_size_ = self.shape_[0]
        if directly_build:
            self._build()
        else:
            self.H_.build_pattern(self.max_leaf_size_, self.eta_)
            super().__init__(float, self.shape_)

    def _build(self):
        if self.built_:
            return
        t0 = time.perf_counter()
        self.H_.build(self.max_leaf_size_, self.eta_, self.eps_aca_)
        self.dev_ = HipHmat(self.H_, precisions=self._precisions, adjoint=self._adjoint,
                            device=self._device)
        self.hmat_time_ = time.perf_counter() - t0
        super().__init__(float, self.shape_)
        self.built_ = True
        if self.verbose_:
            print(f"BEMatrixRectangular {self.kernel_}: {self.shape_[0]} x {self.shape_[1]}, "
                  f"compression ratio {self.getCompression():.5f}, build {self.hmat_time_:.2f} s")

    # ---- LinearOperator ---------------------------------------------------
    def _matvec(self, v: np.ndarray) -> np.ndarray:
        self._build()
        return self.dev_.matvec(np.asarray(v, dtype=np.float64).ravel(), original=True,
                                precision=self.precision)

    def _rmatvec(self, v: np.ndarray) -> np.ndarray:
        self._build()
        return self.dev_.rmatvec(np.asarray(v, dtype=np.float64).ravel(), original=True,
                                 precision=self.precision)

    def matvec_torch(self, x, precision: str | None = None):
        self._build()
        return self.dev_.matvec(x, original=True, precision=precision or self.precision)

    # ---- reference helpers -------------------------------------------------
    def getCompression(self) -> float:
        self._build()
        return self.H_.compression_ratio()

    def getPermutation(self) -> np.ndarray:
        """Source (column) permutation."""
        return self.H_.permutation()

    def getPermutationReceivers(self) -> np.ndarray:
        """Receiver (row) permutation."""
        return self.H_.permutation_receivers()

    def get_omp_threads(self) -> int:
        """Host threads of the H-matrix construction (the matvec runs on the GPU)."""
        n = self.n_openMP_threads_
        return n if n and n > 0 else (os.cpu_count() or 1)

    def getMeshCollocationPoints(self) -> np.ndarray:
        """Collocation points of the receiver mesh, (n_receivers, dim), mesh
        order (BigWhamIO::GetCollocationPoints reads the receiver mesh)."""
        return self.H_.receiver_collocation_points()

    def getCollocationPoints(self) -> np.ndarray:
        """The receivers' collocation points in mesh order.  (The reference
        scatters them through the source permutation, which fits only when
        both meshes have as many points; the receivers are returned as the
        mesh holds them.)"""
        return self.H_.receiver_collocation_points()

    def getSpatialDimension(self) -> int:
        return self.H_.info()[1]

    def _getPattern(self) -> np.ndarray:
        """Rows (i_begin, j_begin, i_end, j_end, FLAG, entry size) as
        BigWhamIO::GetHPattern, rows in receiver points and columns in source
        points: FLAG 0 full rank, 1 low rank."""
        fp, lp = self.H_.fr_pattern(), self.H_.lr_pattern()
        rows = [np.stack([fp[1], fp[2], fp[3], fp[4], np.zeros_like(fp[1]),
                          (fp[3] - fp[1]) * (fp[4] - fp[2])], 1),
                np.stack([lp[1], lp[2], lp[3], lp[4], np.ones_like(lp[1]), lp[5]], 1)]
        return np.concatenate(rows, 0)

    def _getFullBlocks(self, original_order: bool = True) -> csc_matrix:
        """Entries of the full-rank blocks as a sparse (N_receivers, N_sources)
        matrix (bigwham4py.py:668-679)."""
        self._build()
        val, row, col = full_blocks(self.H_, original=original_order)
        return csc_matrix((val, (row, col)), shape=self.shape_)

    def get_max_error_ACA(self) -> float:
        return self.H_.max_aca_error()

    def getMatvecTime(self, N_matvec: int = 100) -> float:
        """Seconds per matvec.  The input has shape[1] entries (the reference
        sizes it with shape[0], which fits only square operators)."""
        x = np.ones(self.shape[1])
        self._matvec(x)
        t0 = time.time()
        for _ in range(N_matvec):
            self._matvec(x)
        return (time.time() - t0) / N_matvec
