import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def device():
    """One device handle for the whole GPU session (one process on the card)."""
    from fastkcna_amd import Device
    dev = Device(0)
    yield dev
    dev.close()


def int_mixture(n, d, centers, seed):
    from fastkcna_amd.datasets import integer_mixture
    return integer_mixture(n, d, centers, seed)


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
