import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "openmm-atmmetaforce-plugin_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def has_gpu():
    try:
        from atmmetaforce import _capi
        return _capi.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    from atmmetaforce import _capi
    _capi.load()  # a missing library fails loudly
    if _capi.device_count() < 1:
        pytest.fail("no GPU visible to libatm_mi355x.so")
    return True
