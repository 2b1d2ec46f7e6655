import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: full-size workloads")


@pytest.fixture(scope="session")
def native_lib():
    from bevy_hair_amd import _native
    return _native.lib()


@pytest.fixture(scope="session")
def renderer():
    from bevy_hair_amd import plugin
    r = plugin.Renderer(0)
    yield r
    r.close()
