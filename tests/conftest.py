import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build_oracle()
    pyoracle.set_ordering(True)
    return pyoracle


@pytest.fixture(scope="session")
def stencils():
    g = np.load(os.path.join(GOLDEN, "type2_stencils.npz"))
    return g["lap"], g["mass"]


@pytest.fixture(scope="session")
def dgx():
    import dendro4_amd
    dendro4_amd.init(0)
    return dendro4_amd
