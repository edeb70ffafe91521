import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmplbm.so on the GPU)")
    config.addinivalue_line("markers", "slow: long CPU test")


def _gpu_count():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return n.value if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 else 0
    except OSError:
        return 0


def pytest_collection_modifyitems(config, items):
    if any("gpu" in it.keywords for it in items) and _gpu_count() == 0:
        skip = pytest.mark.skip(reason="no HIP device visible")
        for it in items:
            if "gpu" in it.keywords:
                it.add_marker(skip)
