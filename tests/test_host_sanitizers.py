"""The host-side C++ (H construction, slab layout, composite, full blocks)
built with g++ -fsanitize=address,undefined and driven through its C-ABI
(tests/asan/host_driver.cpp): any out-of-bounds access, use after free, leak
or undefined behaviour aborts the test program.  CPU only."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "bigwham_dev_amd" / "csrc"
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
         f"-I{ROOT / 'include'}"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_under_sanitizers(tmp_path):
    sources = [ROOT / "tests" / "asan" / "host_driver.cpp",
               *(CSRC / "hip" / n for n in ("layout.cpp", "layout_capi.cpp", "composite.cpp",
                                            "fullblocks.cpp")),
               *sorted((CSRC / "host").glob("hb_*.cpp"))]

    def compile_one(src):
        obj = tmp_path / (src.stem + ".o")
        res = subprocess.run(["g++", *FLAGS, "-c", str(src), "-o", str(obj)],
                             capture_output=True, text=True)
        assert res.returncode == 0, res.stderr[-2000:]
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, sources))
    exe = tmp_path / "host_driver"
    res = subprocess.run(["g++", *FLAGS, *map(str, objs), "-o", str(exe), "-pthread"],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    # verify_asan_link_order=0: the environment may preload a library ahead
    # of the ASan runtime, which ASan otherwise refuses at start-up
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    res = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert res.returncode == 0, (res.stdout[-2000:], res.stderr[-4000:])
    assert "host sanitizer driver ok" in res.stdout
