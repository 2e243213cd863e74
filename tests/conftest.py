import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: full-size workloads")


def _have_gpu() -> bool:
    try:
        from polysolve_amd import linear

        return linear.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _have_gpu():
        pytest.fail("GPU test requested but no GPU / HIP library is available")
    return True


def pytest_collection_modifyitems(config, items):
    # Print the test id before each GPU test starts (the case is named even if a run dies).
    os.environ.setdefault("PYTHONUNBUFFERED", "1")
