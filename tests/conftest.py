import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


def _load_cases(fname):
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", fname)
    with np.load(path, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    cases = {}
    for key, val in data.items():
        name, field = key.split("__")
        cases.setdefault(name, {})[field] = val
    return cases


@pytest.fixture(scope="session")
def fixtures():
    return _load_cases("gsknn_ref_fixtures.npz")


@pytest.fixture(scope="session")
def l1_fixtures():
    return _load_cases("gsknn_ref_l1_fixtures.npz")
