"""Allocation failure on the device (the failure-detection row of the C-ABI's
error contract): with BWHIP_FAIL_ALLOC_AT=k the k-th device or pinned-host
allocation of one C-ABI call fails as if memory were exhausted.  Every
failing create must return BW_ENOMEM with a message, leave no handle and give
every byte back (hipMemGetInfo before and after); a matvec whose per-stream
scratch cannot be allocated fails the same way and the handle keeps working.
The reference exits the process on any CUDA error (hmat_cuda.cc:31-73)."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HipHmat, HostHmat
from bigwham_dev_amd import _native as N
from bigwham_dev_amd.meshes import rect_grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def host():
    h = HostHmat.from_mesh("3DR0-H", *rect_grid(16), [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    return h


def _free_bytes():
    import torch
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info(0)[0]


def test_create_failure_frees_everything(gpu, host, monkeypatch):
    import torch
    lib = N.hip()
    flat = host.flat()
    mask = N.BW_PREC_FP64 | N.BW_PREC_FP32
    # warm-up: code objects, runtime state
    HipHmat(host, precisions=("fp64", "fp32"), adjoint=True).close()
    torch.cuda.synchronize()
    before = _free_bytes()
    n_failed = 0
    for k in range(1, 200):
        monkeypatch.setenv("BWHIP_FAIL_ALLOC_AT", str(k))
        out = C.c_void_p()
        print(f"create with allocation {k} failing", flush=True)
        rc = lib.bw_hip_create(C.byref(flat), mask, N.BW_FLAG_ADJOINT, 0, C.byref(out))
        if rc == N.BW_OK:
            lib.bw_hip_destroy(out)
            break
        msg = lib.bw_hip_last_error().decode()
        assert rc == N.BW_ENOMEM, (k, rc, msg)
        assert "hipMalloc" in msg or "hipHostMalloc" in msg, msg
        assert not out.value
        assert _free_bytes() == before, f"allocation {k}: device memory not given back"
        n_failed += 1
    else:
        pytest.fail("create never succeeded")
    monkeypatch.delenv("BWHIP_FAIL_ALLOC_AT")
    assert n_failed >= 10  # every allocation of the create was made to fail once
    assert _free_bytes() == before


def test_matvec_scratch_failure(gpu, host, monkeypatch):
    """A call from a new stream allocates that stream's scratch; when that
    fails the call returns BW_ENOMEM, frees what it took, and the next call
    succeeds with the right result."""
    import torch
    dev = HipHmat(host, precisions=("fp64",))
    x = np.random.default_rng(2).uniform(-1, 1, host.size)
    xd = torch.from_numpy(x).to(gpu)
    ref = O.matvec_original(host.flat(), x)
    s = torch.cuda.Stream(device=gpu)
    s.wait_stream(torch.cuda.default_stream(gpu))
    y = torch.empty_like(xd)
    before = _free_bytes()
    for k in (1, 2, 3):
        monkeypatch.setenv("BWHIP_FAIL_ALLOC_AT", str(k))
        print(f"matvec on a new stream with allocation {k} failing", flush=True)
        rc = N.hip().bw_hip_matvec(dev._h, C.c_void_p(xd.data_ptr()), C.c_void_p(y.data_ptr()),
                                   N.BW_ORDER_ORIGINAL, N.BW_PREC_FP64, 1,
                                   C.c_void_p(s.cuda_stream))
        assert rc == N.BW_ENOMEM, rc
        assert "scratch" in N.hip().bw_hip_last_error().decode()
        assert _free_bytes() == before
    monkeypatch.delenv("BWHIP_FAIL_ALLOC_AT")
    with torch.cuda.stream(s):
        y = dev.matvec(xd)
    s.synchronize()
    assert np.linalg.norm(y.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)
    assert os.environ.get("BWHIP_FAIL_ALLOC_AT") is None
