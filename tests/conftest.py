import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    pyoracle.build()  # make is a no-op when up to date
    return pyoracle.load()


@pytest.fixture(scope="session")
def kscr_lib():
    from kamisimulator_amd import _lib
    return _lib.load()
