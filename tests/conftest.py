import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "irsounderlbl.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libirs.so")


@pytest.fixture(scope="session")
def ctx():
    from irsounderlbl.api import Context
    c = Context(0)
    yield c
    c.close()
