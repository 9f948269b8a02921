"""pytest configuration: the `gpu` marker and shared fixtures."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library with at least one visible device (no fallback)."""
    from code_mojo_amd import _lib
    lib = _lib.load()
    if lib.bfly_device_count() < 1:
        pytest.fail("gpu test selected but no GPU is visible to libbfly_hip.so")
    return lib
