import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblightgraph on cuda:0)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import lgoracle
    lgoracle.lib()
    return lgoracle


@pytest.fixture(scope="session")
def lga():
    import lightgraph_amd
    lightgraph_amd.lib()
    return lightgraph_amd


@pytest.fixture(scope="session")
def ctx(lga):
    c = lga.Context(0)
    yield c
    c.close()


def rel_err(a, b, floor=1e-6):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b), floor * max(1e-300, float(np.max(np.abs(b))) if b.size else 1.0))
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0
