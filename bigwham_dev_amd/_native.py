"""ctypes bindings of the two native libraries (see include/*.h).

Nothing here falls back to Python: a missing library raises ImportError-like
errors at first use, and every negative status code raises BigWhamError with
the library's own message.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

BW_OK = 0
BW_EINVAL, BW_ENOMEM, BW_EHIP, BW_ESHAPE, BW_ESTATE, BW_ENOTSUP = -1, -2, -3, -4, -5, -6
BW_PREC_FP64 = 1
BW_PREC_FP32 = 2
BW_FLAG_ADJOINT = 1
BW_ORDER_PERMUTED = 0
BW_ORDER_ORIGINAL = 1

_ERRNAMES = {-1: "BW_EINVAL", -2: "BW_ENOMEM", -3: "BW_EHIP", -4: "BW_ESHAPE",
             -5: "BW_ESTATE", -6: "BW_ENOTSUP"}


class BigWhamError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class FlatHmat(C.Structure):
    """Mirror of bw_flat_hmat (include/bigwham_flat.h)."""
    _fields_ = [
        ("dof", C.c_int32), ("reserved", C.c_int32),
        ("n_row_pts", C.c_int64), ("n_col_pts", C.c_int64),
        ("perm0", C.POINTER(C.c_int64)), ("perm1", C.POINTER(C.c_int64)),
        ("n_fr", C.c_int64), ("fr_pattern", C.POINTER(C.c_int64)),
        ("fr_blocks", C.POINTER(C.POINTER(C.c_double))),
        ("n_lr", C.c_int64), ("lr_pattern", C.POINTER(C.c_int64)),
        ("lr_A", C.POINTER(C.POINTER(C.c_double))),
        ("lr_B", C.POINTER(C.POINTER(C.c_double))),
    ]


_host = None
_hip = None

_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_vp = C.c_void_p


def _load(name: str) -> C.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise ImportError(
            f"{path} is not built; run `python -m bigwham_dev_amd.build` "
            "(or __graft_entry__.build())")
    return C.CDLL(str(path))


def host() -> C.CDLL:
    global _host
    if _host is None:
        lib = _load("libbwhost.so")
        lib.bwh_last_error.restype = C.c_char_p
        lib.bwh_create.argtypes = [C.c_char_p, _d, C.c_int64, _i64, C.c_int64, _d, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]
        lib.bwh_create_rect.argtypes = [C.c_char_p, _d, C.c_int64, _i64, C.c_int64, _d,
                                        C.c_int64, _i64, C.c_int64, _d, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.POINTER(_vp)]
        lib.bwh_create_gaussian.argtypes = [C.c_int64, C.c_double, C.c_int, C.c_int,
                                            C.POINTER(_vp)]
        lib.bwh_build_pattern.argtypes = [_vp, C.c_int, C.c_double]
        lib.bwh_build_hmat.argtypes = [_vp, C.c_int, C.c_double, C.c_double]
        lib.bwh_flat.argtypes = [_vp, C.POINTER(FlatHmat)]
        lib.bwh_info.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), _i64]
        lib.bwh_shape.argtypes = [_vp, _i64, _i64]
        lib.bwh_collocation_points.argtypes = [_vp, _d]
        lib.bwh_receiver_collocation_points.argtypes = [_vp, _d]
        lib.bwh_element_normals.argtypes = [_vp, _d]
        lib.bwh_rotation_matrices.argtypes = [_vp, _d]
        lib.bwh_dense_block.argtypes = [_vp, C.c_int64, C.c_int64, C.c_int64, C.c_int64, _d]
        lib.bwh_max_aca_error.argtypes = [_vp, _d]
        lib.bwh_timings.argtypes = [_vp, _d, _d]
        lib.bwh_influence.argtypes = [_vp, C.c_int64, C.c_int64, _d]
        lib.bwh_free.argtypes = [_vp]
        lib.bwh_free.restype = None
        _host = lib
    return _host


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        lib = _load("libbwhip.so")
        lib.bw_hip_last_error.restype = C.c_char_p
        lib.bw_hip_create.argtypes = [C.POINTER(FlatHmat), C.c_int, C.c_int, C.c_int,
                                      C.POINTER(_vp)]
        lib.bw_hip_matvec.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]
        lib.bw_hip_rmatvec.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]
        lib.bw_hip_matvec_noalloc.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, _vp]
        lib.bw_hip_matvec_batched.argtypes = [_vp, C.c_int, _vp, C.c_int64, _vp, C.c_int64,
                                              C.c_int, C.c_int, _vp]
        lib.bw_hip_rmatvec_batched.argtypes = lib.bw_hip_matvec_batched.argtypes
        lib.bw_hip_matvec_batched_strided.argtypes = [_vp, C.c_int, C.c_int, _vp, C.c_int64,
                                                      C.c_int64, _vp, C.c_int64, C.c_int64,
                                                      C.c_int, C.c_int, _vp]
        lib.bw_hip_diagonal.argtypes = [_vp, _d, C.c_int]
        lib.bw_hip_sizes.argtypes = [_vp, _i64, _i64, _i64, _i64, _i64, _i64]
        lib.bw_hip_set_tuning.argtypes = [C.c_int64, C.c_int64]
        lib.bw_hip_set_stacking.argtypes = [C.c_int]
        lib.bw_hip_set_unroll.argtypes = [_vp, C.c_int, C.c_int]
        lib.bw_hip_set_unroll_pass.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
        lib.bw_hip_set_mfma_schedule.argtypes = [_vp, C.c_int, C.c_int]
        lib.bw_hip_set_timing.argtypes = [_vp, C.c_int]
        lib.bw_hip_timing_stats.argtypes = [_vp, _d, _i64]
        lib.bw_hip_destroy.argtypes = [_vp]
        lib.bw_hip_destroy.restype = None
        lib.bw_layout_build.argtypes = [C.POINTER(FlatHmat), C.c_int, C.c_int, C.c_int64,
                                        C.POINTER(_vp)]
        lib.bw_layout_info.argtypes = [_vp, _i64]
        lib.bw_layout_export.argtypes = [_vp, C.c_int, _i64, C.POINTER(C.c_int32), _i64, _d]
        lib.bw_layout_export_rowmap.argtypes = [_vp, C.POINTER(C.c_int32)]
        lib.bw_layout_export_xcd.argtypes = [_vp, C.c_int, C.POINTER(C.c_int32),
                                             C.POINTER(C.c_int32)]
        lib.bw_layout_free.argtypes = [_vp]
        lib.bw_layout_free.restype = None
        lib.bw_layout_last_error.restype = C.c_char_p
        lib.bw_layout_build_multi.argtypes = [C.POINTER(C.POINTER(FlatHmat)), C.c_int, C.c_int,
                                              C.c_int, C.c_int64, C.POINTER(_vp)]
        lib.bw_hip_create_multi.argtypes = [C.POINTER(C.POINTER(FlatHmat)), C.c_int, C.c_int,
                                            C.c_int, C.c_int, C.POINTER(_vp), _i64, _i64]
        lib.bw_hip_graph_create.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(_vp)]
        lib.bw_hip_graph_create_batched.argtypes = [_vp, C.c_int, C.c_int, _vp, C.c_int64,
                                                    C.c_int64, _vp, C.c_int64, C.c_int64,
                                                    C.c_int, C.c_int, C.POINTER(_vp)]
        lib.bw_hip_stream_release.argtypes = [_vp, _vp]
        lib.bw_hip_set_fp32_rank_tolerance.argtypes = [C.c_double]
        lib.bw_hip_entries.argtypes = [_vp, C.c_int, _i64, _i64, _i64, _i64]
        lib.bw_flat_recompress.argtypes = [C.POINTER(FlatHmat), C.c_double, C.POINTER(_vp)]
        lib.bw_flat_recompressed_view.argtypes = [_vp]
        lib.bw_flat_recompressed_view.restype = C.POINTER(FlatHmat)
        lib.bw_flat_recompressed_stats.argtypes = [_vp, _i64, _i64, _i64]
        lib.bw_flat_recompress_last_error.restype = C.c_char_p
        lib.bw_flat_recompressed_free.argtypes = [_vp]
        lib.bw_flat_recompressed_free.restype = None
        lib.bw_hip_graph_launch.argtypes = [_vp, _vp]
        lib.bw_hip_graph_destroy.argtypes = [_vp]
        lib.bw_hip_graph_destroy.restype = None
        lib.bw_flat_compose.argtypes = [C.POINTER(C.POINTER(FlatHmat)), C.c_int, C.POINTER(_vp)]
        lib.bw_flat_composite_view.argtypes = [_vp]
        lib.bw_flat_composite_view.restype = C.POINTER(FlatHmat)
        lib.bw_flat_composite_offsets.argtypes = [_vp, _i64, _i64]
        lib.bw_flat_compose_last_error.restype = C.c_char_p
        lib.bw_flat_composite_free.argtypes = [_vp]
        lib.bw_flat_composite_free.restype = None
        lib.bw_hip_create_selection.argtypes = [C.POINTER(FlatHmat), _i64, C.c_int64, _i64,
                                                C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                                                C.POINTER(_vp)]
        lib.bw_flat_full_blocks.argtypes = [C.POINTER(FlatHmat), C.c_int, _i64, C.c_int64, _i64,
                                            C.c_int64, _i64, _d, _i64, _i64]
        lib.bw_flat_full_blocks_last_error.restype = C.c_char_p
        _hip = lib
    return _hip


def check_host(rc: int) -> None:
    if rc != BW_OK:
        raise BigWhamError(rc, host().bwh_last_error().decode())


def check_hip(rc: int) -> None:
    if rc != BW_OK:
        raise BigWhamError(rc, hip().bw_hip_last_error().decode())


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_d)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_i64)


# exported symbols each header declares (checked by the CPU tests)
HOST_SYMBOLS = ["bwh_create", "bwh_create_rect", "bwh_create_gaussian", "bwh_build_pattern",
                "bwh_build_hmat", "bwh_flat", "bwh_info", "bwh_shape", "bwh_collocation_points",
                "bwh_receiver_collocation_points", "bwh_element_normals",
                "bwh_rotation_matrices", "bwh_dense_block", "bwh_max_aca_error",
                "bwh_timings", "bwh_free", "bwh_last_error", "bwh_influence"]
HIP_SYMBOLS = ["bw_hip_create", "bw_hip_matvec", "bw_hip_matvec_noalloc", "bw_hip_rmatvec", "bw_hip_matvec_batched",
               "bw_hip_rmatvec_batched", "bw_hip_matvec_batched_strided",
               "bw_hip_diagonal", "bw_hip_sizes", "bw_hip_set_tuning", "bw_hip_set_stacking", "bw_hip_set_unroll", "bw_hip_set_unroll_pass", "bw_hip_set_mfma_schedule",
               "bw_hip_set_timing",
               "bw_hip_timing_stats", "bw_hip_destroy", "bw_hip_last_error",
               "bw_layout_build", "bw_layout_info", "bw_layout_export", "bw_layout_free",
               "bw_layout_export_rowmap", "bw_layout_export_xcd",
               "bw_layout_last_error", "bw_layout_build_multi", "bw_hip_create_multi",
               "bw_hip_graph_create", "bw_hip_graph_launch", "bw_hip_graph_destroy",
               "bw_hip_graph_create_batched", "bw_hip_stream_release",
               "bw_hip_set_fp32_rank_tolerance", "bw_hip_entries", "bw_flat_recompress",
               "bw_flat_recompressed_view", "bw_flat_recompressed_stats",
               "bw_flat_recompress_last_error", "bw_flat_recompressed_free",
               "bw_flat_compose", "bw_flat_composite_view", "bw_flat_composite_offsets",
               "bw_flat_compose_last_error", "bw_flat_composite_free",
               "bw_hip_create_selection", "bw_flat_full_blocks",
               "bw_flat_full_blocks_last_error"]
