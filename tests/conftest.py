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


@pytest.fixture(autouse=True)
def no_sort_faults(request):
    """After every test that ran on the device: no sort gave up a look-back wait (a sort that did
    would have left a wrong order)."""
    yield
    if request.node.get_closest_marker("gpu") is not None:
        import dynamis_amd
        assert dynamis_amd.sort_faults(clear=True) == 0, "a sort's look-back wait gave up"


@pytest.fixture(params=["oracle", pytest.param("device", marks=pytest.mark.gpu)])
def engine(request):
    """A factory of worlds over encoded scenes: the oracle's or the device's."""
    if request.param == "oracle":
        oracle = request.getfixturevalue("oracle")
        return lambda scene, capacity=None: oracle.world(scene, capacity)
    request.getfixturevalue("device")
    from dynamis_amd.world import World

    def make(scene, capacity=None):
        world = World()
        world.create(scene, capacity)
        return world
    return make
