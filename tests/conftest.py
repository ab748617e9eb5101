import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X; calls libmqw's HIP kernels on device 0")


@pytest.fixture(scope="session")
def ctx():
    import multistageqw_amd as M
    c = M.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def c1_recs():
    import numpy as np
    return np.fromfile(os.path.join(GOLDEN, "sk_n10_m2_seed1.bin"), "<f8").reshape(64, 55)
