import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with a visible device; fails (never skips) otherwise so a
    GPU run cannot pass without the native path."""
    import rtac_simulation_amd as R
    n = R.device_count()
    assert n >= 1, "gpu tests need a visible HIP device"
    return R
