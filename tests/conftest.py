import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtsunami_hip.so on the GPU)")


@pytest.fixture(scope="session")
def repo():
    return REPO


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as orc
    orc.lib()
    return orc


@pytest.fixture(scope="session")
def tsu():
    import tsunami_pact_amd
    tsunami_pact_amd.library()
    return tsunami_pact_amd


@pytest.fixture(scope="session")
def aligner(tsu):
    a = tsu.Aligner(4, 6, 2)
    yield a
    a.close()
