import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


def golden(name):
    import numpy

    return numpy.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def hip():
    """The HIP package with a usable device; fails loudly when the GPU test
    runs without a device or without the built library."""
    import blackhorizon_amd as bh
    from blackhorizon_amd import _lib

    _lib.require_device()
    return bh
