"""The C-ABI libraries load, export every symbol include/*.h declares, and
reject bad input with status codes (no GPU needed: every case here fails
before any device call)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from bigwham_dev_amd import BigWhamError, HostHmat
from bigwham_dev_amd import _native as N
from bigwham_dev_amd.meshes import segment_line

ROOT = Path(__file__).resolve().parents[1]


def _declared(header):
    text = (ROOT / "include" / header).read_text()
    return set(re.findall(r"^\s*(?:int|void|const char\*|const \w+\s*\*)\s+(\w+)\s*\(", text, re.M))


def test_headers_match_exports():
    host = N.host()
    hip = N.hip()
    for name in _declared("bigwham_host.h"):
        assert hasattr(host, name), name
    for name in _declared("bigwham_hip.h"):
        assert hasattr(hip, name), name
    assert _declared("bigwham_host.h") == set(N.HOST_SYMBOLS)
    assert _declared("bigwham_hip.h") == set(N.HIP_SYMBOLS)


def test_host_errors():
    with pytest.raises(BigWhamError) as e:
        HostHmat.from_mesh("9XX-H", np.zeros((2, 2)), np.array([[0, 1]]), [1.0, 0.25])
    assert e.value.code == -1
    with pytest.raises(BigWhamError) as e:
        HostHmat.from_mesh("2DS0-H", np.zeros((2, 2)), np.array([[0, 5]]), [1.0, 0.25])
    assert e.value.code == -4
    with pytest.raises(BigWhamError):
        HostHmat.from_mesh("2DS0-H", *segment_line(4), [1.0, 0.7])  # nu > 0.5
    h = HostHmat.from_mesh("2DS0-H", *segment_line(10), [1.0, 0.25])
    with pytest.raises(BigWhamError) as e:
        h.flat()  # nothing built yet
    assert e.value.code == -5
    with pytest.raises(BigWhamError):
        h.dense_block(0, 0, 11, 1)


def _bad_flat(h, mutate):
    f = h.flat()
    g = N.FlatHmat()
    C.pointer(g)[0] = f
    return mutate(g) or g


def test_engine_rejects_bad_layout_before_device():
    h = HostHmat.from_mesh("2DS0-H", *segment_line(200), [1.0, 0.25])
    h.build(32, 3.0, 1e-6)
    hip = N.hip()
    out = C.c_void_p()
    f = h.flat()
    # precision mask / flags
    assert hip.bw_hip_create(C.byref(f), 0, 0, -1, C.byref(out)) in (-1, -3)
    # a pattern that does not tile the matrix
    g = N.FlatHmat()
    C.memmove(C.byref(g), C.byref(f), C.sizeof(f))
    g.n_fr = f.n_fr - 1
    assert hip.bw_hip_create(C.byref(g), 1, 0, -1, C.byref(out)) == -4
    # a broken permutation
    bad = np.zeros(f.n_col_pts, dtype=np.int64)
    C.memmove(C.byref(g), C.byref(f), C.sizeof(f))
    g.perm1 = bad.ctypes.data_as(C.POINTER(C.c_int64))
    assert hip.bw_hip_create(C.byref(g), 1, 0, -1, C.byref(out)) == -4
    # null handle / vectors
    assert hip.bw_hip_matvec(None, None, None, 0, 1, 1, None) == -1
    assert hip.bw_hip_diagonal(None, None, 0) == -1
    assert hip.bw_hip_set_tuning(-1, 0) == -1
    lay = C.c_void_p()
    assert hip.bw_layout_build(C.byref(g), 1, 0, 0, C.byref(lay)) == -4
