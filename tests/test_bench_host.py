"""The host legs of bench.py, on the CPU: the CPU baseline (the oracle's
OpenMP restatement of Hmat::matvecOriginal, src/hmat/hmatrix/hmat.cc:572-635,
built for this host) and the host path's GMRES (scipy's GMRES with the Jacobi
preconditioner of bigwham4py.py:373-380 on that matvec), and the one-GPU
contract of the command line."""
import subprocess
import sys

import numpy as np
import pytest

import bench
import oracle as O
from bigwham_dev_amd import solvers


@pytest.fixture(scope="module")
def c1():
    cfg = bench.CONFIGS["c1"]
    coor, conn = bench.make_mesh(*cfg["mesh"])
    return bench._build_host(cfg["kernel"], coor, conn, cfg)


def test_cpu_baseline_and_host_gmres(c1):
    import torch
    n = c1.size
    x = np.random.default_rng(5).uniform(-1, 1, n)
    y, host_matvec, cb = bench.cpu_baseline(c1, x, budget_s=0.5)
    flat = c1.flat()
    ref = O.matvec_original(flat, x)
    assert np.linalg.norm(y - ref) <= 1e-14 * np.linalg.norm(ref)
    assert np.linalg.norm(host_matvec(x) - ref) <= 1e-13 * np.linalg.norm(ref)
    assert cb["kind"] == "port" and cb["value"] > 0 and cb["cores"] >= 1
    d = O.diagonal(flat, original=True)
    b = np.zeros(n)
    b[c1.dof - 1::c1.dof] = 1.0
    out = bench.host_gmres(host_matvec, d, b, 1e-8, 30)
    assert out["scipy_info"] == 0 and out["rel_residual_fp64"] <= 1e-8
    # the device solver's algorithm on CPU tensors needs about as many
    # matvecs (both restart every 30 iterations, preconditioned the same way)
    inv = torch.from_numpy(1.0 / d)
    _, its, _ = solvers.gmres(lambda v: torch.from_numpy(host_matvec(v.numpy())),
                              torch.from_numpy(b), lambda v: inv * v, tol=1e-8, restart=30)
    assert abs(out["matvecs"] - its) <= 0.1 * its + 30


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(bench.ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
