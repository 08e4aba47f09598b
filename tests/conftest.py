import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def renderer():
    from vork_amd import native
    return native.Renderer(0)


try:  # load torch's HIP runtime before the library's (one runtime per process)
    import torch  # noqa: F401
except ImportError:
    pass
