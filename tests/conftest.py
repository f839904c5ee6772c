import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def fclib():
    import foil_cfd_amd
    return foil_cfd_amd.load()


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with a visible device; a GPU test without one fails."""
    import foil_cfd_amd
    lib = foil_cfd_amd.load()
    n = lib.fc_device_count()
    assert n >= 1, "no HIP device visible to a gpu-marked test"
    return lib
