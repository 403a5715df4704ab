import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def ctx():
    import spec_spars_amd as ss
    c = ss.Context(0)
    yield c
    c.close()


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def components(n, u, v):
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    A = sp.coo_matrix((np.ones(len(u)), (np.asarray(u), np.asarray(v))), shape=(n, n))
    return cg.connected_components(A, directed=False)[0]
