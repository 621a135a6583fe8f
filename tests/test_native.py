"""Sanitizer runs of the host restatement of the tile binning (tests/native/bin_restatement.cpp).

The restatement follows the binning kernels' index arithmetic (keys, segments, cursor places,
the owner-computes deposit, the fused kernel's row-block halo search, and the tile scatter's
neighbourhood bins, reserved ranges, far counts and bin codes after a move of every particle)
with std::thread workers, and checks its invariants against a direct CIC deposit.  It is built twice with g++:
with AddressSanitizer + UndefinedBehaviorSanitizer (out-of-bounds indices, overflow) and with
ThreadSanitizer (races between workers: two tiles writing one cell would be one).
"""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "native" / "bin_restatement.cpp"

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")


@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_binning_restatement_under_sanitizer(tmp_path, san):
    exe = tmp_path / f"bins_{san.split(',')[0]}"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", f"-fsanitize={san}",
           "-fno-omit-frame-pointer", "-fno-sanitize-recover=all", str(SRC), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0",
           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
           "TSAN_OPTIONS": "halt_on_error=1", "PATH": "/usr/bin:/bin"}
    p = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "FAILED" not in p.stdout and p.stdout.count(": ok") == 12
