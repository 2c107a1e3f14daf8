import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs through the C-ABI")


@pytest.fixture(scope="session")
def orc():
    """The FP32 CPU oracle (ctypes wrappers)."""
    from tests import oracle_api

    return oracle_api.Oracle(fp64=False)


@pytest.fixture(scope="session")
def orc64():
    """The FP64 CPU oracle (ctypes wrappers)."""
    from tests import oracle_api

    return oracle_api.Oracle(fp64=True)
