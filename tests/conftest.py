import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def ctx():
    """HIP context on device 0.  GPU tests never fall back: no device = fail."""
    import muontomography_amd as M
    c = M.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def small_set():
    from muontomography_amd import workloads as W
    g, rays, info = W.muons("small", n_muons=20_000)
    return g, rays, info


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    s = np.maximum(np.abs(b), 1e-300)
    return d / s
