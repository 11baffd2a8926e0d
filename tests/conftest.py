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


def ucell_from_box(box):
    """Unit cell (rows = cell vectors) of an orthorhombic box (lengths, 90 deg
    angles), as Box::SetupFromXyzAbg gives it for that case."""
    box = np.asarray(box, np.float64)
    assert np.allclose(box[3:], 90.0), box
    return np.diag(box[:3])


def min_image_distance(a, b, ucell):
    """Minimum-image distance between points a and b in the periodic cell
    `ucell` (rows = cell vectors): the fractional difference wrapped to the
    nearest image, then the 27 neighbouring images searched (exact for any
    cell whose reduced vectors are no longer than its diagonal; pinned against
    the compiled DIST2_ImageNonOrtho in test_oracle.py)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    inv = np.linalg.inv(ucell)
    f = (b - a) @ inv
    f -= np.round(f)
    best = np.inf
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                d = (f + np.array([i, j, k], np.float64)) @ ucell
                best = min(best, float(d @ d))
    return np.sqrt(best)
