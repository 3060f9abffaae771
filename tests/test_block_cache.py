"""Host checks of the device block cache: best fit, splitting, coalescing back to
whole segments, release, the out-of-memory fallback and single-thread ownership,
over a fake segment source (tests/block_cache_check.cpp), built with g++ and the
address and undefined-behaviour sanitizers."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_block_cache_host_checks(tmp_path):
    exe = tmp_path / "block_cache_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", "-fsanitize=address,undefined",
                    "-I", os.path.join(ROOT, "horkxg_amd", "csrc", "hip"),
                    os.path.join(ROOT, "tests", "block_cache_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "block cache checks passed" in out.stdout
