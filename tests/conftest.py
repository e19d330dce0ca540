import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libnnpe.so")


@pytest.fixture(scope="session")
def water_c0():
    from nnpengine_amd import system as S
    return S.water_box(256, 19.73, seed=1)


@pytest.fixture(scope="session")
def water_c1():
    from nnpengine_amd import system as S
    return S.water_box(8000, 62.1, seed=7)


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).mean() / max((b ** 2).mean(), 1e-300)))
