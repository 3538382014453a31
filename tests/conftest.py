import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libhobak_hip.so")
    config.addinivalue_line("markers", "slow: long CPU-side run")


def golden(name):
    import numpy as np

    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name} not generated")
    return np.load(path, allow_pickle=False)


@pytest.fixture(scope="session")
def gpu_ctx():
    from hobakv1_amd import lib

    ctx = lib.Context(0)
    yield ctx
    ctx.close()
