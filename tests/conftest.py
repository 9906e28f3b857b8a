import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs a ROCm GPU (MI355X) and the built libsdegpu.so')


@pytest.fixture(scope='session')
def gpu():
    """The torch module, after checking that a GPU and the library exist."""
    from sdepy_amd import _lib
    return _lib.require_gpu()
