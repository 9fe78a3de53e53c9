import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950)")


@pytest.fixture(scope="session")
def port():
    from oracle.oracle import Oracle

    return Oracle("port")


@pytest.fixture(scope="session")
def ref():
    from oracle.oracle import LIBS, Oracle

    if not os.path.exists(LIBS["reference"]):
        pytest.skip("reference CPU build (oracle/_ref) not present")
    return Oracle("reference")


@pytest.fixture(scope="session")
def core():
    import gaffa_amd

    return gaffa_amd._core


@pytest.fixture(scope="session")
def gpu(core):
    if core.gpu_device_count() == 0:
        pytest.fail("GPU test requested but no HIP device is visible")
    return core
