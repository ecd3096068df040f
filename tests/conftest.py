import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: larger parity cases")


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(HERE, "golden", "reference_test_data.npz"))


@pytest.fixture(scope="session")
def ctx():
    import raptor_amd as ra
    c = ra.Context(0)
    yield c
    c.close()
