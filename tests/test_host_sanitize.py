"""The library's host C++ (csrc/dgk_host.cpp) under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/host_sanitize_driver.cpp is built with it
by g++ (no GPU code, no HIP) and run over every operator file, several
meshes and the error paths, with buffers of exactly the documented sizes."""
import os
import shutil
import subprocess

import pytest

TESTS = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(TESTS)
CSRC = os.path.join(REPO, "scale-dg-kernel-extraction_amd", "csrc")


def test_host_code_under_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "host_sanitize")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           os.path.join(TESTS, "host_sanitize_driver.cpp"), os.path.join(CSRC, "dgk_host.cpp"),
           "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, os.path.join(REPO, "operator_data")], capture_output=True,
                       text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert "host checks passed" in r.stdout
    assert "Sanitizer" not in r.stderr, r.stderr[-6000:]
