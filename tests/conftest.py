import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


def read_layout(path):
    """XcT_init free-format reader (fortran/advect_fibers.f90:195,318-321)."""
    tok = open(path).read().split()
    n = int(tok[0])
    vals = np.array(tok[1:1 + 6 * n], dtype=np.float64).reshape(n, 6)
    return vals[:, :3].ravel().copy(), vals[:, 3:].ravel().copy()


@pytest.fixture(scope="session")
def layout100():
    return read_layout(os.path.join(GOLDEN, "inputs", "XcT_init100.in"))


def load_ref(variant, step, name):
    return np.loadtxt(os.path.join(GOLDEN, "reference_100", "100_" + variant, "%d_%s.ref" % (step, name)))


VARIANTS = ["analytic_direct", "analytic_gmres", "numeric_direct", "numeric_gmres"]
QUANTITIES = [("BVec", "b"), ("XVec", "x"), ("TRANSVel", "vel"), ("ROTVel", "rot"),
              ("POS", "pos"), ("ORIENT", "orient")]
