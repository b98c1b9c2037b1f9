"""Host-side code built with AddressSanitizer and UndefinedBehaviorSanitizer (gcc, CPU only).

The C oracle is driven through every entry point on small inputs, coincident points included
(tests/sanitize_oracle.c), and the soft binary64 quantiser check (tests/softfp_check.cpp) runs
on a reduced case count; any out-of-bounds access, leak or undefined operation aborts the run.
"""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _need(tool):
    if shutil.which(tool) is None:
        pytest.skip(f"{tool} not available")


def test_oracle_under_asan_ubsan():
    _need("gcc")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "san_oracle")
        subprocess.run(["gcc", "-std=c11", *SAN, "-ffp-contract=off",
                        os.path.join(ROOT, "tests", "sanitize_oracle.c"),
                        os.path.join(ROOT, "oracle", "bh_oracle.c"), "-o", exe, "-lm",
                        "-lpthread"], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=300)
        assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr[-4000:]


def test_softfp_check_under_asan_ubsan():
    _need("g++")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "softfp_san")
        subprocess.run(["g++", "-std=c++17", *SAN, "-ffp-contract=off",
                        os.path.join(ROOT, "tests", "softfp_check.cpp"), "-o", exe], check=True)
        r = subprocess.run([exe, "200000"], capture_output=True, text=True, env=ENV, timeout=300)
        assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr[-4000:]
