"""Writes tests/golden/ref_pattern_fixtures.npz from the compiled reference
harness oracle/_ref/libbwref.so (built from the reference's own clustering,
pattern and segment-kernel sources by oracle/ref/build_ref.sh).

Stored per mesh and clustering variant: the collocation points given to the
reference, its permutation and its FRB / LRB pattern tables (rows 0-4);
plus We_segment_0 outputs on 20 seeded receivers.  tests/test_golden.py checks
the host library of bigwham_dev_amd against them, so the pins hold where the reference
sources are not mounted.  Run: python tests/golden/make_ref_fixtures.py"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from bigwham_dev_amd import HostHmat  # noqa: E402
from bigwham_dev_amd.meshes import penny_delaunay, rect_grid, segment_line  # noqa: E402

REF = ROOT / "oracle" / "_ref" / "libbwref.so"
OUT = Path(__file__).resolve().parent / "ref_pattern_fixtures.npz"

MESHES = {"2DS0-H": (lambda: segment_line(2000), 32),
          "3DR0-H": (lambda: rect_grid(60), 64),
          "3DT0-H": (lambda: penny_delaunay(3000, seed=5), 64)}


def main() -> None:
    ref = C.CDLL(str(REF))
    D, I = C.POINTER(C.c_double), C.POINTER(C.c_int64)
    ref.ref_pattern.argtypes = [D, C.c_int64, C.c_int, C.c_int64, C.c_double, C.c_int, C.c_int,
                                C.c_int64, I, I, I, I, I]
    ref.ref_we_segment_0.argtypes = [C.c_double] * 5 + [D]
    ref.ref_we_segment_0.restype = None
    out = {}
    for kernel, (mesh, leaf) in MESHES.items():
        coor, conn = mesh()
        for hom in (0, 1):
            h = HostHmat.from_mesh(kernel, coor, conn, [1.0, 0.25], homogeneous_size_pattern=bool(hom))
            h.build_pattern(leaf, 3.0)
            pts = np.ascontiguousarray(h.collocation_points())
            n, dim = pts.shape
            cap = 200000
            fr = np.zeros(6 * cap, np.int64)
            lr = np.zeros(6 * cap, np.int64)
            nfr, nlr = C.c_int64(), C.c_int64()
            perm = np.zeros(n, np.int64)
            ref.ref_pattern(pts.ctypes.data_as(D), n, dim, leaf, 3.0, hom, -1, cap,
                            fr.ctypes.data_as(I), C.byref(nfr), lr.ctypes.data_as(I),
                            C.byref(nlr), perm.ctypes.data_as(I))
            key = f"{kernel}_h{hom}"
            out[key + "_points"] = pts
            out[key + "_perm"] = perm
            out[key + "_fr"] = fr[: 6 * nfr.value].reshape(-1, 6).T[:5].copy()
            out[key + "_lr"] = lr[: 6 * nlr.value].reshape(-1, 6).T[:5].copy()
    rng = np.random.default_rng(4)
    args = np.column_stack([rng.uniform(0.1, 2.0, 20), np.full(20, 1.3 / 2.42),
                            np.full(20, 0.21), rng.uniform(-2, 2, 20), rng.uniform(-2, 2, 20)])
    W = np.zeros((20, 6))
    for i, a in enumerate(args):
        ref.ref_we_segment_0(*map(float, a), W[i].ctypes.data_as(D))
    out["we_segment_0_args"] = args
    out["we_segment_0_values"] = W
    np.savez_compressed(OUT, **out)
    print(OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
