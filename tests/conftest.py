import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "simplex-benchmarks_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def smf():
    import simplexmf
    simplexmf.lib()
    return simplexmf


@pytest.fixture(scope="session")
def orc():
    import oracle_py
    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="session")
def gpu(smf):
    """The HIP runtime with at least one device; GPU tests fail loudly without it."""
    n = smf.device_count()
    if n < 1:
        pytest.fail("gpu test selected but no HIP device is visible")
    return smf


def rel_err(a, b):
    import numpy as np
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
