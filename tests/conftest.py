import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def chl():
    from plant_amd import chl as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as m
    m.lib()
    return m
