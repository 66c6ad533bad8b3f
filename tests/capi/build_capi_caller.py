"""Builds tests/capi/_build/capi_caller, the C++ caller of the C-ABI run by
tests/test_gpu_capi.py.  Host code only (g++ with the HIP runtime API
header), linked against the in-tree libbwhip.so / libbwhost.so with an
$ORIGIN-relative rpath so the binary follows the tree."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
OUT = HERE / "_build" / "capi_caller"


def build() -> Path:
    OUT.parent.mkdir(exist_ok=True)
    lib = ROOT / "bigwham_dev_amd" / "lib"
    rel = os.path.relpath(lib, OUT.parent)
    tmp = OUT.with_suffix(".tmp")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
           f"-I{ROOT / 'include'}", f"-I{ROCM / 'include'}", str(HERE / "capi_caller.cpp"),
           "-o", str(tmp), f"-L{lib}", "-lbwhip", "-lbwhost", f"-L{ROCM / 'lib'}", "-lamdhip64",
           f"-Wl,-rpath,$ORIGIN/{rel}", f"-Wl,-rpath,{ROCM / 'lib'}", "-pthread"]
    print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("capi_caller build failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
