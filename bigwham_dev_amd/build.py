"""Builds the native libraries of bigwham_dev_amd in-tree.

  lib/libbwhost.so   host H-matrix construction (g++, std::thread)
  lib/libbwhip.so    MI355X engine: HIP kernels for gfx950 + C-ABI (hipcc)

Run as ``python -m bigwham_dev_amd.build``.  The oracle under ``oracle/`` is
built by ``oracle/build_oracle.py``, which is test infrastructure.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB = PKG / "lib"
CSRC = PKG / "csrc"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HOST_SOURCES = sorted((CSRC / "host").glob("hb_*.cpp"))
# libbwhip.so: the kernels and their C-ABI (hipcc), plus host-only C++ (the
# slab layout, composition, full blocks, recompression) compiled by g++ with
# the host ISA of libbwhost.so
HIP_SOURCES = [CSRC / "hip" / "engine.hip"]
HIP_HOST_SOURCES = [CSRC / "hip" / "layout.cpp", CSRC / "hip" / "layout_capi.cpp",
                    CSRC / "hip" / "composite.cpp", CSRC / "hip" / "fullblocks.cpp",
                    CSRC / "hip" / "recompress.cpp"]
HOST_FLAGS = ["-std=c++17", "-O3", "-march=x86-64-v3", "-fPIC", "-Wall", "-Wextra"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    res = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {res.returncode})")


def build_host() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libbwhost.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", *HOST_FLAGS, "-shared", "-o", tmp, *HOST_SOURCES, "-pthread"])
    os.replace(tmp, out)
    return out


def build_hip() -> Path:
    LIB.mkdir(exist_ok=True)
    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    out = LIB / "libbwhip.so"
    tmp = out.with_suffix(".so.tmp")
    objs = []
    for src in HIP_HOST_SOURCES:
        obj = obj_dir / (src.stem + ".o")
        _run(["g++", *HOST_FLAGS, "-c", "-o", obj, src])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = obj_dir / (src.stem + ".o")
        _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-c",
              "-Wno-unused-value", "-Wno-unused-result",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-o", obj, src])
        objs.append(obj)
    _run([HIPCC, f"--offload-arch={ARCH}", "-fPIC", "-shared", "-o", tmp, *objs, "-pthread"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_host()
    build_hip()


if __name__ == "__main__":
    build_all()
