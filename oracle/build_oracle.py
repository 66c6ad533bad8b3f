"""Builds the CPU oracle (test infrastructure): oracle/lib/liboracle.so.

The oracle restates the reference's CPU matvec path in C (hmat_oracle.c).
When /root/reference is present, ref/build_ref.sh also compiles the few
reference sources that build standalone into oracle/_ref/libbwref.so.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build_oracle() -> Path:
    out_dir = HERE / "lib"
    out_dir.mkdir(exist_ok=True)
    out = out_dir / "liboracle.so"
    tmp = out.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-o", str(tmp),
           str(HERE / "hmat_oracle.c"), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, out)
    return out


NATIVE_FLAGS = ["-O3", "-march=native", "-fopenmp"]


def build_oracle_native(out_dir) -> Path:
    """The same oracle built for the CPU it runs on (-O3 -march=native), for
    the host-core baseline timing; built on the machine that times it."""
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    out = out_dir / "liboracle_native.so"
    cmd = ["gcc", *NATIVE_FLAGS, "-std=c11", "-fPIC", "-shared", "-o", str(out),
           str(HERE / "hmat_oracle.c"), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("native oracle build failed")
    return out


def build_ref() -> Path | None:
    if not Path("/root/reference").is_dir():
        return None
    script = HERE / "ref" / "build_ref.sh"
    res = subprocess.run(["bash", str(script)], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("reference harness build failed")
    return HERE / "_ref" / "libbwref.so"


if __name__ == "__main__":
    print(build_oracle())
    print(build_ref())
