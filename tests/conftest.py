import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnuray.so on the device)")


@pytest.fixture(scope="session")
def nr():
    """The product binding with device 0 initialised."""
    import nuray_amd

    nuray_amd.init(0)
    return nuray_amd


@pytest.fixture(scope="session")
def po():
    import pyoracle

    return pyoracle


@pytest.fixture(scope="session")
def scenes():
    from nuray_amd import scenes as s

    return s


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)
