import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) and libdachshund_gpu.so")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def fixture_pixels():
    """tests/small_pixel_data.bin of the reference: 2000 Pixel records."""
    return np.fromfile(os.path.join(GOLDEN, "small_pixel_data.bin"),
                       dtype="<f8").reshape(-1, 5)


@pytest.fixture(scope="session")
def ora():
    from oracle import ora as o
    o.lib()
    return o
