import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def kzdev():
    """Initialise device 0 once per session and return the package."""
    import openkhuzdulgraph_amd as kz

    kz.init(0)
    return kz
