"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every entry
point of refcpu.cpp, for every problem family, method, error controller and
nonlinear-solver stage mask, and with invalid inputs (oracle/sanitize_main.cpp,
built by `make -C oracle sanitize`).  Any out-of-bounds access, leak or undefined
arithmetic aborts the driver."""
import os
import pathlib
import shutil
import subprocess

import pytest

ORACLE = pathlib.Path(__file__).resolve().parents[1] / "oracle"


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_oracle_clean_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", str(ORACLE), "sanitize"], check=True, capture_output=True, timeout=300)
    # verify_asan_link_order=0: the process environment may preload other libraries
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(ORACLE / "_build" / "sanitize_refcpu")], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "OK: 0 failures" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
