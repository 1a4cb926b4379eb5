import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libbsmps.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def golden_m8():
    import numpy as np
    return dict(np.load(os.path.join(GOLDEN, "m8_site_golden.npz")))


@pytest.fixture(scope="session")
def golden_haf():
    import numpy as np
    return dict(np.load(os.path.join(GOLDEN, "hafnian_general.npz")))


def same_bits(a, b):
    """Exact equality of complex/real arrays with == semantics (+0 == -0), NaN-free."""
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    return bool(np.all(a == b))
