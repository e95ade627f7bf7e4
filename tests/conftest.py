import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ref():
    return np.load(os.path.join(GOLDEN, "ref_fixtures.npz"), allow_pickle=False)


@pytest.fixture(scope="session")
def tables():
    return np.load(os.path.join(GOLDEN, "green_tables.npz"), allow_pickle=False)


def wrapped(x):
    """Difference folded into (-pi, pi]."""
    return (np.asarray(x) + np.pi) % (2 * np.pi) - np.pi
