import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def toy7():
    import detray_amd as da
    det = da.Detector.toy(4, 7)
    yield det


@pytest.fixture(scope="session")
def toy3():
    import detray_amd as da
    det = da.Detector.toy(4, 3)
    yield det


@pytest.fixture(scope="session")
def telescope():
    import detray_amd as da
    det = da.Detector.telescope([100.0 * i for i in range(1, 11)], 500.0, 500.0, 0.15, 60.0)
    yield det


@pytest.fixture(scope="session")
def solenoid():
    import detray_amd as da
    yield da.MagneticField.solenoid()
