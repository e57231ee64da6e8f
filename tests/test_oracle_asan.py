"""The oracle under AddressSanitizer, LeakSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle asan` builds oracle/asan_check.cpp with honfft_oracle.cpp
instrumented; the driver runs every plan kind (D = 1..4, float and double,
the three constructors) through every oracle entry point on small inputs.  Any
out-of-bounds access, leak or undefined operation makes it exit non-zero.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_is_clean_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_asan", "oracle_asan")], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "plan cases ok" in r.stdout
