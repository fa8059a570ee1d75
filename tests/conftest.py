import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

_GPU = None


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


def gpu_available() -> bool:
    """Ask the HIP runtime through libonionpir.so itself (no torch in the process)."""
    global _GPU
    if _GPU is None:
        try:
            from onionpirv2_amd import _capi
            _GPU = _capi.device_count() > 0
        except ImportError:
            if os.path.exists("/dev/kfd"):
                raise  # a GPU machine without the built library: fail loudly
            _GPU = False
    return _GPU


def pytest_runtest_setup(item):
    if "gpu" in item.keywords and not gpu_available():
        pytest.skip("no GPU visible")
