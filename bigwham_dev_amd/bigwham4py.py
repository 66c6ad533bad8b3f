"""BEMatrix: the Python face of the reference (bigwham4py.BEMatrix,
interfaces/python/bigwham4py/bigwham4py.py:49-485) on the MI355X engine.

Same constructor arguments, same LinearOperator semantics (shape (N, N),
float64, matvec in the original DoF ordering), same helper methods for the
matvec path: getCompression, getPermutation, getCollocationPoints,
getSpatialDimension, H_diag, H_jacobi_prec, get_diagonal, getMatvecTime,
get_max_error_ACA, _getPattern.  Construction stays on the host (libbwhost);
the matvec runs on the GPU (libbwhip).  `useCuda` is accepted for signature
compatibility; there is no CPU matvec fallback in this package.

Extras of this engine: `precision="fp32"` for FP32-factor matvecs,
`matvec_torch(x)` on device tensors, `rmatvec` (H^T), and `solve()`
(FP64 or mixed-precision GMRES on the device).
"""
from __future__ import annotations

import time

import numpy as np
from scipy.sparse import diags
from scipy.sparse.linalg import LinearOperator

from .hmat import HipHmat, HostHmat

kernels_id = ["2DS0-H", "3DT0-H", "3DR0-H", "3DR0-H-mode1"]


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

    def get_max_error_ACA(self) -> float:
        return self.H_.max_aca_error()

    def get_diagonal(self) -> np.ndarray:
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
