import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: minutes of CPU work")


def gpu_available() -> bool:
    try:
        import simd_bigint_amd as sba
        return sba.lib().sba_gpu_device_count() > 0
    except Exception:
        return False
