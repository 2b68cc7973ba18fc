import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def rel_l2(a, b):
    """||a - b|| / ||b|| over all elements (complex allowed)."""
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.oracle.LIB_PATH):
        oracle.oracle.build()
    oracle.set_threads(min(16, os.cpu_count() or 1))
    return oracle.solvers()
