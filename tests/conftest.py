import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def gpu_ready():
    """The HIP library loaded and at least one device visible; fails (not
    skips) otherwise, so a GPU run cannot pass on a missing extension."""
    from eggpu_amd import native

    native.lib()
    count = native.device_count()
    assert count >= 1, "no HIP device visible to libegx.so"
    return count


@pytest.fixture
def gpu_env(monkeypatch):
    monkeypatch.setenv("EASYGRAPH_ENABLE_GPU", "TRUE")
    monkeypatch.setenv("EASYGRAPH_GPU_STRICT_ERRORS", "TRUE")
