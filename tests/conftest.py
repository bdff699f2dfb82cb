import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    if not os.path.exists(pyoracle.LIB):
        pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def gpu_lib():
    import spaqr_public_amd as sq
    return sq.load_library()


@pytest.fixture(scope="session")
def ktest():
    import ctypes as C
    path = os.path.join(ROOT, "spaqr_public_amd", "_lib", "libspaqr_ktest.so")
    if not os.path.exists(path):
        raise RuntimeError("kernel test library not built; run __graft_entry__.build()")
    return C.CDLL(path)
