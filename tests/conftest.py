import importlib
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmpm.so")


@pytest.fixture(scope="session")
def mpm():
    return importlib.import_module("water-sim-cuda_amd")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o

    if not o.LIB_PATH.exists():
        o.build()
    return o


@pytest.fixture(scope="session")
def gpu(mpm):
    """Skips GPU tests when no HIP device is visible (CPU container); on a GPU
    box the engine must load and run, there is no fallback."""
    if not mpm.gpu_available():
        pytest.skip("no HIP device visible")
    if not mpm.LIB_PATH.exists():
        pytest.fail("libmpm.so is not built")
    return mpm
