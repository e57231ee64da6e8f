"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/cpp/abi_asan is the library's sources (host pass instrumented,
gadgetron_amd/csrc/Makefile `asan`) linked with a driver that calls every C ABI
family with invalid and null arguments, valid descriptors on a host without a
GPU, and from two threads (the thread-local error message).  Any stray access,
leak or undefined operation on those host paths makes it exit non-zero.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "abi_asan")


def test_host_code_is_clean_under_sanitizers():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "gadgetron_amd", "csrc"), "asan"], check=True)
    try:
        import gadgetron_amd as ga
        if ga.device_count() > 0:
            pytest.skip("runs on a host without a GPU (the device paths are covered by the gpu tests)")
    except ImportError:
        pass
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([BIN], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "checks ok" in r.stdout
