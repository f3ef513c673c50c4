import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_ctypes as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def lib_cmpmc():
    import cmpmc_amd
    cmpmc_amd.capi.load()
    return cmpmc_amd


@pytest.fixture(scope="session")
def gold():
    """The golden vectors written by the reference's own build (tests/golden/make_golden.py)."""
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "reference_vectors.npz"))
