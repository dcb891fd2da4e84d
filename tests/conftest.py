import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through libvrt.so)")
    config.addinivalue_line("markers", "slow: long-running GPU case at full benchmark size")


@pytest.fixture(scope="session")
def vrt():
    import __graft_entry__ as ge
    mod = ge.load_package()
    if not os.path.exists(mod.LIB_PATH):
        ge.build_lib()
    mod.load_library()
    return mod


@pytest.fixture(scope="session")
def orc():
    import oracle_ctypes
    oracle_ctypes.lib()
    return oracle_ctypes


@pytest.fixture(scope="session")
def gpu(vrt):
    if vrt.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return vrt
