"""bigwham_dev_amd — MI355X-native hierarchical-matrix matvec engine for BigWham.

Host construction (mesh, kernels, cluster tree, ACA) lives in libbwhost.so;
the matvec path (H x, H^T x, batched, FP64 / FP32 factors) runs as
hand-written HIP kernels for gfx950 in libbwhip.so.  See include/*.h for the
C-ABI.
"""
from ._native import BigWhamError  # noqa: F401
from .hmat import HipHmat, HostHmat, set_tuning  # noqa: F401

__all__ = ["BigWhamError", "HipHmat", "HostHmat", "set_tuning"]
