import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT), str(ROOT / "geos-chem_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


@pytest.fixture(scope="session")
def fullchem():
    from gckpp_amd import mechanism
    return mechanism.load("fullchem")


@pytest.fixture(scope="session")
def hg():
    from gckpp_amd import mechanism
    return mechanism.load("Hg")


@pytest.fixture(scope="session")
def mech_of():
    from gckpp_amd import mechanism
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = mechanism.load(name)
        return cache[name]
    return get
