import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu():
    """Skips the test when no GPU is visible (the CPU container); on a GPU host it runs."""
    import fastload_amd as fl
    if fl.device_count() < 1:
        pytest.skip("no GPU visible")
    return fl
