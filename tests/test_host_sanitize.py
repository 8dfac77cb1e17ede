"""The library's host-side validation and packing (csrc/bvp_host.cpp, the code behind
every entry point's argument checks) under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/native/check_host.cpp feeds it valid batches (and
checks the packed arrays) and every kind of invalid instance and option, expecting
the documented error codes."""
import os
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_validation_and_packing_under_asan_ubsan(tmp_path):
    exe = tmp_path / "check_host"
    subprocess.run(["g++", "-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-o", str(exe), str(ROOT / "tests" / "native" / "check_host.cpp"),
                    str(ROOT / "boundaryvaluediffeq.jl_amd" / "csrc" / "bvp_host.cpp")],
                   check=True, capture_output=True, timeout=300)
    # verify_asan_link_order=0: the process environment may preload other libraries
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "OK: 0 failures" in r.stdout
