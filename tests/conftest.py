import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def params137():
    from openifs_amd import cloudsc as C

    return C.default_params(900.0, 137)


@pytest.fixture(scope="session")
def params91():
    from openifs_amd import cloudsc as C

    return C.default_params(900.0, 91)
