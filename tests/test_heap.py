"""The server's job-buffer heap (raypier_optics_amd/csrc/rpx_heap.h), built
over malloc with g++ and AddressSanitizer and driven through random
allocate/free sequences (tests/native/heap_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "raypier_optics_amd", "csrc")


def test_block_heap_random_sequences(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "heap_check")
    src = os.path.join(HERE, "native", "heap_check.cpp")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-I", CSRC, src, "-o", exe]
    build = subprocess.run(cmd, capture_output=True, text=True)
    if build.returncode != 0 and "sanitize" in build.stderr:
        cmd = [c for c in cmd if not c.startswith("-fsanitize")]
        build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([exe, "6"], capture_output=True, text=True, timeout=240)
    if run.returncode != 0 and "does not come first" in run.stderr:
        # a library preloaded into every process keeps the sanitizer's
        # runtime from loading first: check without the sanitizers
        cmd = [c for c in cmd if not c.startswith("-fsanitize")]
        build = subprocess.run(cmd, capture_output=True, text=True)
        assert build.returncode == 0, build.stderr
        run = subprocess.run([exe, "6"], capture_output=True, text=True, timeout=240)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "heap ok: 6 seeds" in run.stdout
