"""Shared pytest configuration: the ``gpu`` marker and import paths."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libwalrus_hip.so")


@pytest.fixture
def rng():
    return np.random.default_rng(137)


def random_symmetric(rng, n, complex_=True, scale=1.0):
    """Symmetric (B + B^T)/2 with iid normal entries (complex: Re, Im each N(0, scale^2))."""
    B = rng.normal(size=(n, n)) * scale
    if complex_:
        B = B + 1j * rng.normal(size=(n, n)) * scale
    return (B + B.T) / 2


def relerr(a, b):
    return abs(complex(a) - complex(b)) / max(abs(complex(b)), 1e-300)
