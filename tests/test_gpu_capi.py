"""The C-ABI driven from C++ with no Python in the loop
(tests/capi/capi_caller.cpp, built by __graft_entry__.build()): host build,
one upload, the host low-rank factors released (as the Hmat adapter does),
then the facade's post-build calls (nbOfEntries, full blocks, diagonal),
host- and device-pointer matvecs, FP32, H^T, three right-hand sides through
the batched entry point, and the error codes.
Its outputs are compared here with the oracle on the same H (rebuilt from
the same mesh by the same host library).  Bars: FP64 1e-12, FP32 1e-5."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HostHmat
from bigwham_dev_amd.meshes import rect_grid, segment_line

pytestmark = pytest.mark.gpu

CALLER = Path(__file__).resolve().parent / "capi" / "_build" / "capi_caller"


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("kernel,mesh", [("3DR0-H", lambda: rect_grid(24)),
                                         ("2DS0-H", lambda: segment_line(1500))])
def test_cpp_caller(gpu, tmp_path, kernel, mesh):
    if not CALLER.exists():  # a test program, built with the libraries by build()
        pytest.skip(f"{CALLER} is not built: run __graft_entry__.build()")
    coor, conn = mesh()
    coor = np.ascontiguousarray(coor, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    with open(tmp_path / "mesh.bin", "wb") as f:
        np.array([coor.shape[0], coor.shape[1], conn.shape[0], conn.shape[1]], np.int64).tofile(f)
        coor.tofile(f)
        conn.tofile(f)
    print(f"capi caller {kernel} nodes={coor.shape[0]}", flush=True)
    res = subprocess.run([str(CALLER), kernel, str(tmp_path / "mesh.bin"), str(tmp_path)],
                         capture_output=True, text=True, timeout=300)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0 and "capi_caller: ok" in res.stdout

    h = HostHmat.from_mesh(kernel, coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    f = h.flat()
    n = h.size
    rd = lambda name: np.fromfile(tmp_path / f"{name}.bin", dtype=np.float64)
    x = rd("x")
    assert x.size == n
    y_ref = O.matvec_original(f, x)
    y64 = rd("y64")
    assert rel(y64, y_ref) <= 1e-12
    assert rel(rd("y32"), y_ref) <= 1e-5
    assert rel(rd("yt"), O.rmatvec_original(f, x)) <= 1e-12
    assert np.array_equal(rd("diag"), O.diagonal(f, original=True))
    entries, nfb = np.fromfile(tmp_path / "entries.bin", dtype=np.int64)
    assert entries == O.nb_entries(f) == h.nb_entries()
    fp = h.fr_pattern()
    assert nfb == int(np.sum(h.dof * (fp[3] - fp[1]) * h.dof * (fp[4] - fp[2])))
    from bigwham_dev_amd.hmat import full_blocks
    vals, ri, ci = full_blocks(h, original=True)
    assert np.array_equal(rd("fb_val"), vals)
    assert np.array_equal(np.fromfile(tmp_path / "fb_row.bin", dtype=np.int64), ri)
    assert np.array_equal(np.fromfile(tmp_path / "fb_col.bin", dtype=np.int64), ci)
    Yb = rd("ybatch").reshape(3, n).T
    for j, c in enumerate((1.0, -1.0, 2.0)):
        assert rel(Yb[:, j], c * y_ref) <= 1e-12, j
