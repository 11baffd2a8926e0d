"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
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
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reflib():
    from oracle.oracle import RefLib, ref_available
    if not ref_available():
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    return RefLib()


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name + ".npz"))
    return load


@pytest.fixture(scope="session")
def ctx():
    import torch
    import cpptraj_amd
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    c = cpptraj_amd.Context(0)
    yield c
    c.synchronize()
    c.close()


def tri_to_square(tri, F):
    """Packed upper triangle (Matrix<float> TRIANGLE) -> symmetric (F, F)."""
    M = np.zeros((F, F), dtype=np.float64)
    iu = np.triu_indices(F, 1)
    M[iu] = tri
    M[(iu[1], iu[0])] = tri
    return M
