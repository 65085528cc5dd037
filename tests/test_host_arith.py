"""Host checks of the exact-arithmetic shortcuts in the HIP kernels (division by the odd mean
window via multiply and FMA correction; integer square root of the gradient from a float
estimate; the u16 range test of the y smoothing as bounds on the unrounded value):
tests/cpp/arith_check.c restates them with the same operations and compares them with the
plain IEEE division, the reference's double sqrt + truncation and roundf."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_exact_arithmetic_shortcuts():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "arith_check")
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", os.path.join(HERE, "cpp", "arith_check.c"), "-o", exe, "-lm"],
                       check=True)
        # every 97th float bit pattern for each window (all exponents and signs are covered)
        r = subprocess.run([exe, "97"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    lines = r.stdout.splitlines()
    assert len(lines) == 3 and all(" 0 mismatches" in l for l in lines), r.stdout
