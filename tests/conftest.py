import ctypes
import functools
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP solver)")
    config.addinivalue_line("markers", "slow: long CPU test")


@functools.lru_cache(maxsize=None)
def hip_device_count() -> int:
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        try:
            hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        except OSError:
            return 0
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


@pytest.fixture
def gpu():
    """GPU tests skip (never pass) without a device, like the reference's exit code 77."""
    if hip_device_count() < 1:
        pytest.skip("no HIP device")
    return 0
