import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_present() -> bool:
    if os.environ.get("BGK_FORCE_NO_GPU"):
        return False
    try:
        from bgk3d_gpu_amd import runtime as rt
        return rt.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_present():
        pytest.fail("GPU test selected but no usable GPU / HIP library (build first)")
    return True
