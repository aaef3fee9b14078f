import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmzhip.so")


@pytest.fixture(scope="session")
def mzo():
    import mzo as m
    m.build()
    return m


@pytest.fixture(scope="session")
def hip():
    """The product package (HIP path); GPU tests only."""
    import __graft_entry__ as ge
    return ge.load_package()
