import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def wavecar():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "wavecar_ga4.npz")))


@pytest.fixture(scope="session")
def wavecar_ncl():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "wavecar_ncl.npz")))
