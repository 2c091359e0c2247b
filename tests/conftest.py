import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the GPU path through the C ABI")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    if not os.path.exists(O._LIB_PATH):
        O.build()
    O.set_threads(min(8, os.cpu_count() or 1))
    return O


def rel_err(a, b):
    """max-norm error of a relative to the max-norm of b"""
    import numpy as np
    den = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(a - b).max()) / den
