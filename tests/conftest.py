import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _has_gpu():
    try:
        import bmfr_amd
        return bmfr_amd.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Fails (not skips) when a gpu-marked test runs without a usable GPU or
    without the built library: the HIP path must never be silently absent."""
    import bmfr_amd
    n = bmfr_amd.device_count()
    assert n > 0, "no HIP device visible"
    return n


@pytest.fixture(scope="session")
def oracle():
    from oracle import bind
    bind.load()
    return bind
