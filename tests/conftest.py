import gzip
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def espp_positions():
    """tests/LennardJones/positions.txt of the reference (32,768 atoms, box 33.8585^3)"""
    with gzip.open(os.path.join(GOLDEN, "espp_positions.txt.gz"), "rt") as f:
        return np.loadtxt(f).reshape(-1, 3)
