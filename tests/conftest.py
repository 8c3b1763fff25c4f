import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "adjoint-optimization_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU oracle run (tens of seconds)")


@pytest.fixture(scope="session")
def golden():
    import json

    import numpy as np

    d = dict(np.load(os.path.join(ROOT, "tests", "golden", "ref256.npz")))
    with open(os.path.join(ROOT, "tests", "golden", "ref256.json")) as f:
        meta = json.load(f)
    return d, meta
