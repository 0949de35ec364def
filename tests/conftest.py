import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def engine():
    # A missing or broken libidash_hip.so is an error here, never a skip.
    from idash2019_2_amd import Engine, _abi
    n = ctypes.c_int()
    _abi.check(_abi.lib().ihe_device_count(ctypes.byref(n)))
    if n.value == 0:
        pytest.skip("no GPU visible")
    e = Engine(0)
    yield e
    e.close()
