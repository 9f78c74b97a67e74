import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X GPU (runs through libtenes_amd.so)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One library context on cuda:0 for the whole GPU test session."""
    import tenes_amd

    ctx = tenes_amd.Context(0, seed=1234)
    yield ctx
    ctx.close()
