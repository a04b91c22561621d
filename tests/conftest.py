import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the GPU)")


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="session")
def kq():
    """The product library (fails loudly if it is not built)."""
    import kuqcd_amd.api as A
    A.L.load()
    return A


@pytest.fixture
def gpu(kq):
    """A fresh library state on the GPU for one test."""
    kq.set_verbosity(0)
    assert kq.device_count() >= 1, "no HIP device visible"
    assert kq.kuqcd_init() == kq.KUQCD_SUCCESS
    yield kq
    kq.kuqcd_gauge_free()
    kq.kuqcd_end()
