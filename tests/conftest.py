import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


def _device_count() -> int:
    from spac_amd import abi

    try:
        return int(abi.load().spac_device_count())
    except Exception:
        return 0


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: a gpu-marked test on a machine without a device is an error, not a skip."""
    n = _device_count()
    if n < 1:
        pytest.fail("gpu test selected but no HIP device is visible")
    return n
