import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="session")
def cap0():
    from cuvhs_amd import native
    return native.generate(0)


@pytest.fixture(scope="session")
def ref0(oracle, cap0):
    return oracle.decode(cap0, 28.0, 64, trace=True)


@pytest.fixture(scope="session")
def cap1_prefix():
    """First 140 fields of the 600-field 28 MHz workload."""
    from cuvhs_amd import native
    return native.generate(1, n_samples=140 * 468140)


@pytest.fixture(scope="session")
def ref1_prefix(oracle, cap1_prefix):
    return oracle.decode(cap1_prefix, 28.0, 64, trace=True)


def s16(cap, n):
    """RawReader conversion of a u8 chunk, zero-filled to n samples."""
    raw = np.zeros(n)
    m = min(n, cap.size)
    raw[:m] = (cap[:m].astype(np.float64) - 128.0) * 256.0
    return raw
