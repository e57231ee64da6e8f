import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgtnufft.so on the device)")


def _gpu_count():
    try:
        from gadgetron_amd import device_count
        return device_count(), ""
    except Exception as e:  # reported by the fixture
        return 0, f"{type(e).__name__}: {e}"


@pytest.fixture(scope="session")
def gpu():
    n, why = _gpu_count()
    if n == 0:
        pytest.fail(f"no GPU visible to libgtnufft.so; gpu tests must run on an MI355X ({why or 'count 0'})")
    return n
