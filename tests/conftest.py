import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def engine():
    import knightking_amd as kk
    eng = kk.Engine()
    yield eng
    eng.close()


@pytest.fixture(scope="session")
def karate():
    """dataset/karate.txt of the reference: 34 vertices, both directions listed, weights in [1,5)."""
    import numpy as np
    d = np.loadtxt(os.path.join(ROOT, "tests", "golden", "karate.txt"))
    return d[:, 0].astype(np.uint32), d[:, 1].astype(np.uint32), d[:, 2].astype(np.float32)
