"""The integer emulation of the reference's FP64 Morton arithmetic (csrc/bh_softfp.h), built
for the host and checked against native binary64 on random and boundary cases."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_softfp_matches_binary64():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "softfp_check")
        subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off",
                        os.path.join(ROOT, "tests", "softfp_check.cpp"), "-o", exe], check=True)
        r = subprocess.run([exe, "1500000"], capture_output=True, text=True)
        assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
