import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def oracle():
    from tests.oracle_binding import load
    return load()


@pytest.fixture(scope="session")
def device():
    """The HIP device; a GPU test fails (never skips) when none is usable."""
    import dynamis_amd
    n = dynamis_amd.device_count()
    assert n > 0, "no HIP device visible: GPU tests need an MI355X"
    dynamis_amd._native.call("dyn_set_device", 0)
    return 0
