"""Shared test setup: import paths, the `gpu` marker and common helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def tol(n):
    """Stated FP32 bar: relative L2 error <= 1e-6 * log2(N) against FP64."""
    return 1e-6 * max(np.log2(max(n, 2)), 1.0)


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN)
