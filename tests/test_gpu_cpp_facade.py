"""The C++ facade (include/apr_amd.hpp) used the way reference code calls
APRConverter / ParticleData / APRTreeNumerics / APRFilter / APRNumerics /
APRReconstruction, checked bit-for-bit against the CPU oracle
(tests/cpp/facade_test.cpp, built by __graft_entry__.build())."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_test")


def test_cpp_facade():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/facade_test is not built (run __graft_entry__.build())")
    print("running tests/cpp/facade_test", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "facade ok" in r.stdout
