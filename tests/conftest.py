import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs an MI355X (HIP device) and the built library')


def _has_gpu():
    try:
        from imate_amd import _lib
        _lib.device_properties()
        return True
    except Exception:
        return False


@pytest.fixture(scope='session')
def gpu_lib():
    from imate_amd import _lib
    lib = _lib.load()
    props = _lib.device_properties()
    assert props['num_devices'] >= 1
    return lib
