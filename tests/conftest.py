import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtlr on the device)")


@pytest.fixture(scope="session")
def tlr():
    return ge.load_package()


@pytest.fixture(scope="session")
def orc():
    return ge.load_oracle()
