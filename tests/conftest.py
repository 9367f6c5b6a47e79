import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libns3d_hip.so")


def rel_l2(a, b):
    """Relative L2 distance ||a-b|| / ||b|| over a list of arrays (complex or real)."""
    if not isinstance(a, (list, tuple)):
        a, b = [a], [b]
    num = sum(float(np.sum(np.abs(np.asarray(x) - np.asarray(y)) ** 2)) for x, y in zip(a, b))
    den = sum(float(np.sum(np.abs(np.asarray(y)) ** 2)) for y in b)
    return (num / den) ** 0.5 if den > 0 else num ** 0.5


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build()
    return oracle
