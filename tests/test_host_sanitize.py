"""The host C++ (world generator, schedule, preprocessing, input and restart files,
calibration search) and the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer: `make -C watergap2_amd/csrc sanitize` builds
tests/host_sanitize.cpp with them (g++/gcc, no HIP code), and any report fails
the run (-fno-sanitize-recover)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_is_clean_under_asan_and_ubsan():
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("g++/gcc not installed")
    subprocess.run(["make", "-C", os.path.join(ROOT, "watergap2_amd", "csrc"), "sanitize"], check=True,
                   capture_output=True, text=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "_build", "host_sanitize")], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "host_sanitize ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
