import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"fixture {name} missing")
    return np.load(path)


def maxrel(a, b):
    """max |a-b| / max |b| (normwise relative error); NaN entries of b are
    "not compared", and a must be finite wherever b is."""
    a, b = np.asarray(a), np.asarray(b)
    m = np.isfinite(b)
    assert np.isfinite(a[m]).all(), "non-finite values where the reference is finite"
    scale = np.max(np.abs(b[m])) if m.any() else 0.0
    err = np.max(np.abs(a[m] - b[m])) if m.any() else 0.0
    return err / scale if scale > 0 else err
