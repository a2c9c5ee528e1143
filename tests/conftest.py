import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    return pyoracle.Oracle(omp=False)


@pytest.fixture(scope="session")
def oracle_omp():
    import pyoracle

    return pyoracle.Oracle(omp=True)
