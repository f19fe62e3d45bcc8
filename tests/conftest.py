import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def interior(a, prm):
    """interior view of a halo-included field (numpy index = Fortran index + 1)"""
    return a[2:2 + prm.nx, 2:2 + prm.ny, 2:2 + prm.nz]


def rel_l2(a, ref):
    ref = np.asarray(ref)
    d = float(np.linalg.norm(np.asarray(a) - ref))
    n = float(np.linalg.norm(ref))
    return d / n if n > 0 else d


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.load_oracle()
    return oracle


@pytest.fixture(scope="session")
def reference_mod():
    import oracle
    if oracle.load_reference() is None:
        pytest.skip("reference Fortran build (oracle/_ref) not present")
    return oracle
