import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger sizes")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    lib = oracle_lib.load("oracle")
    if lib is None:
        oracle_lib.build(ref=os.path.isdir("/root/reference/src"))
        lib = oracle_lib.load("oracle")
    return oracle_lib


@pytest.fixture(scope="session")
def ctx():
    import libapr_amd as A
    return A.Context(0, keep_intermediates=True)
