import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


def _np(x):
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.asarray(x)


def rel(a, b):
    """Relative 2-norm difference ||a - b|| / ||b|| (numpy or torch)."""
    a, b = _np(a), _np(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def tensor_block_tol(om, lx, ly):
    """Bound on the difference of two FP64 codes' real-space tensor block at
    offset (lx, ly): 64 eps |S|, with S the oracle's weight-error scale
    (tensor_element_abssum: the change of every entry per weight error of one
    unit of each weight vector's largest value; it bounds the sum of |terms|
    as well).  Each term carries the rounding of its weight and of the
    layered kernel (tens of operations in the recursions and exponentials).
    Against blocks built with the reference's quad-precision weights, the
    reference's own FP64 blocks and the oracle's stay within 2.2 eps |S| at
    the fixture offsets; the GPU's and the oracle's reach 38 eps |S| at the
    far offset (175, 0) of a 176x224 grid."""
    return 64 * np.finfo(np.float64).eps * np.linalg.norm(om.tensor_element_abssum(lx, ly))


@pytest.fixture(scope="session")
def golden():
    path = os.path.join(ROOT, "tests", "golden", "reference_vectors.npz")
    return np.load(path, allow_pickle=False)


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle as po
    po.lib()
    return po
