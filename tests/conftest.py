import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def ctx():
    import renzoku_amd as rz
    return rz.Ctx(0)


@pytest.fixture(scope="session")
def cornell():
    from renzoku_amd import scenes
    return scenes.cornell_box(256, 256)


@pytest.fixture(scope="session")
def small_room():
    from renzoku_amd import scenes
    return scenes.room(128, 128, target_tris=20_000)


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
