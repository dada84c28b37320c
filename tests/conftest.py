import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.npz")
REF_DATA = os.path.join(ROOT, "tests", "golden", "ref_data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN)


def dense_from_golden(g, name, m):
    if f"{name}/K" in g:
        return g[f"{name}/K"]
    K = np.zeros((m, m))
    K[g[f"{name}/K_row"], g[f"{name}/K_col"]] = g[f"{name}/K_val"]
    return K


def rel_err(a, b):
    """||a - b|| / ||b|| (0 when both vanish)."""
    nb = np.linalg.norm(b)
    na = np.linalg.norm(a - b)
    return na / nb if nb > 0 else na
