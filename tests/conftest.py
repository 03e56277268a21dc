import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: longer CPU-side test")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def daliborka():
    from uncertaintyte_amd import jacob
    return jacob.read_jacob(os.path.join(GOLDEN, "04_Daliborka.jacob"))


@pytest.fixture(scope="session")
def cube():
    from uncertaintyte_amd import jacob
    return jacob.read_jacob(os.path.join(GOLDEN, "01_Cube.jacob"))


@pytest.fixture(scope="session")
def daliborka_oracle(daliborka):
    import oracle
    return oracle.pipeline(daliborka, debug=True)
