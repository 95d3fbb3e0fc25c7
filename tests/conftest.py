import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "fest-3d_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    config.addinivalue_line("markers", "ref: needs the amdflang build of the reference modules (oracle/_ref)")


def relerr(a, b):
    """max |a-b| / max |b| over finite entries; inf if the finite masks differ."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    fa, fb = np.isfinite(a), np.isfinite(b)
    if (fa != fb).any():
        return np.inf
    if not fa.any():
        return 0.0
    scale = np.abs(b[fa]).max()
    d = np.abs(a[fa] - b[fa]).max()
    return d / scale if scale > 0 else d


@pytest.fixture
def rel():
    return relerr
