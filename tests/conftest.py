import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built librender_hip.so")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def renderer():
    from vtrace_amd import renderer as R
    r = R.get_renderer()
    yield r


@pytest.fixture
def fresh_renderer():
    """A renderer with no volumes or instances (cleanup + entry)."""
    from vtrace_amd import renderer as R
    r = R.reset_renderer()
    yield r
