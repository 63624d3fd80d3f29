import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP path)')


def rel_err(a, b):
    """Norm-wise relative error ||a - b|| / ||b|| (0 when both vanish)."""
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    d = np.linalg.norm((a - b).ravel())
    if nb == 0.0:
        return d
    return d / nb


def phys(e, nr, nph):
    """The physical part of a PPP array (rows < nr, packed columns < np/2)."""
    return np.asarray(e)[:nr, :nph, :]


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
