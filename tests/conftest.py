"""Shared test setup: import paths, the ``gpu`` marker, common fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "lcs.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture
def rng():
    return np.random.default_rng(0)


def rand_halo(rng, dims, h, dtype=np.float64):
    """A halo array filled with uniform [0,1) values everywhere, ghosts included
    (the reference's per-phase tests call Random.rand! on whole arrays)."""
    shape = tuple(n + 2 * h for n in dims)
    return np.asfortranarray(rng.random(shape).astype(dtype))


def interior(A, h):
    return A[h:-h, h:-h, h:-h]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0)
