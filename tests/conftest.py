import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long statistical run")


@pytest.fixture(scope="session")
def goldens():
    import json

    import numpy as np
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    arr = np.load(os.path.join(here, "reference_goldens.npz"))
    with open(os.path.join(here, "reference_scalars.json")) as f:
        sc = json.load(f)
    return arr, sc
