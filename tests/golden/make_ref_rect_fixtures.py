"""Writes tests/golden/ref_pattern_rect_fixtures.npz from the compiled
reference harness oracle/_ref/libbwref.so (ref_pattern_rect: the reference's
own cluster() of each point set and hmatrixTreeIxJ + createPattern,
hierarchical_representation.h:88-137, cluster.cc:213-374).

Stored per source / receiver mesh pair of tests/rect_meshes.py: both
collocation point sets given to the reference; and per clustering variant its
two permutations and its FRB / LRB pattern tables (rows 0-4).
tests/test_rect_host.py checks the host library against them, so the pins
hold where the reference sources are not mounted.
Run: python tests/golden/make_ref_rect_fixtures.py"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))
from bigwham_dev_amd import HostHmat  # noqa: E402
from rect_meshes import KERNELS, pair  # noqa: E402

REF = ROOT / "oracle" / "_ref" / "libbwref.so"
OUT = Path(__file__).resolve().parent / "ref_pattern_rect_fixtures.npz"


def reference_pattern(ref, p0, p1, leaf, eta, hom):
    D, I = C.POINTER(C.c_double), C.POINTER(C.c_int64)
    ref.ref_pattern_rect.argtypes = [D, C.c_int64, D, C.c_int64, C.c_int, C.c_int64, C.c_double,
                                     C.c_int, C.c_int64, I, I, I, I, I, I]
    p0 = np.ascontiguousarray(p0)
    p1 = np.ascontiguousarray(p1)
    cap = 400000
    fr = np.zeros(6 * cap, np.int64)
    lr = np.zeros(6 * cap, np.int64)
    nfr, nlr = C.c_int64(), C.c_int64()
    q0 = np.zeros(len(p0), np.int64)
    q1 = np.zeros(len(p1), np.int64)
    ref.ref_pattern_rect(p0.ctypes.data_as(D), len(p0), p1.ctypes.data_as(D), len(p1),
                         p0.shape[1], leaf, eta, hom, cap, fr.ctypes.data_as(I), C.byref(nfr),
                         lr.ctypes.data_as(I), C.byref(nlr), q0.ctypes.data_as(I),
                         q1.ctypes.data_as(I))
    return (q0, q1, fr[: 6 * nfr.value].reshape(-1, 6).T[:5].copy(),
            lr[: 6 * nlr.value].reshape(-1, 6).T[:5].copy())


def main() -> None:
    ref = C.CDLL(str(REF))
    out = {}
    for kernel in KERNELS:
        cs, ks, cr, kr, leaf = pair(kernel)
        for hom in (0, 1):
            h = HostHmat.from_meshes(kernel, cs, ks, cr, kr, [1.0, 0.25],
                                     homogeneous_size_pattern=bool(hom))
            p0, p1 = h.receiver_collocation_points(), h.collocation_points()
            q0, q1, fr, lr = reference_pattern(ref, p0, p1, leaf, 3.0, hom)
            key = f"{kernel}_h{hom}"
            out[kernel + "_rec_points"] = p0
            out[kernel + "_src_points"] = p1
            out[key + "_perm0"] = q0
            out[key + "_perm1"] = q1
            out[key + "_fr"] = fr
            out[key + "_lr"] = lr
    np.savez_compressed(OUT, **out)
    print(OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
