import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def seeded_psi(shape, grids, seed=1, phase=False, noise=0.01):
    """Gaussian exp(-r^2/2)(1 + noise U[-1,1]) with an optional random phase, on x-fastest grids."""
    rng = np.random.default_rng(seed)
    mesh = np.meshgrid(*grids[::-1], indexing="ij")
    r2 = sum(m ** 2 for m in mesh)
    psi = np.exp(-r2 / 2) * (1 + noise * rng.uniform(-1, 1, shape))
    if phase:
        psi = psi * np.exp(1j * rng.uniform(0, 2 * np.pi, shape))
    return psi.astype(np.complex128)
