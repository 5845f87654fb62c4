import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls through the C ABI")


@pytest.fixture(scope="session")
def ctx():
    import senk_amd as S

    n = S.device_count()
    if n == 0:
        pytest.fail("GPU test selected but no HIP device is visible")
    c = S.Context(0)
    yield c
    c.close()
