"""The reference's own known-answer tests, run through the HIP path."""
import numpy as np
import pytest

import gpue_amd as G
from tests.conftest import seeded_psi

pytestmark = pytest.mark.gpu


def test_parsum_kat():
    # src/unit_test.cu:841-947: 2+2i on 64^2 (and 16^3) with dx = dy = dz = 0.1
    # becomes 2/sqrt(32768 dx dy (dz)) exactly
    with G.Context(dimnum=2, xDim=64, yDim=64, xMax=3.2, yMax=3.2) as c:
        c.upload(np.full((64, 64), 2 + 2j))
        c.parsum()
        w = c.download()
        assert np.all(w.real == 2 / np.sqrt(32768.0 * 0.1 * 0.1))
        assert np.all(w.imag == 2 / np.sqrt(32768.0 * 0.1 * 0.1))
    with G.Context(dimnum=3, xDim=16, yDim=16, zDim=16, xMax=0.8, yMax=0.8, zMax=0.8) as c:
        c.upload(np.full((16, 16, 16), 2 + 2j))
        c.parsum()
        w = c.download()
        assert np.all(w.real == 2 / np.sqrt(32768.0 * 0.1 * 0.1 * 0.1))


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_evolve_ho_energy(dim):
    # evolve_test (src/unit_test.cu:1162-1278): the harmonic ground state has E = d/2
    # (hbar omega units) within 1 %, and real-time evolution keeps E within 1 %.
    n = 32
    with G.Context(dimnum=dim, xDim=n, yDim=n, zDim=n, xMax=6, yMax=6, zMax=6, gdt=1e-3, dt=1e-3,
                   corotating=1, omega=0.0) as c:
        grids = [c.grid(a) for a in range(dim)]
        c.upload(seeded_psi(c.shape, grids, seed=dim))
        c.parsum()
        r, en = c.evolve(30001, 0, energy_every=2000, thr=1e-4)
        e, _ = c.energy()
        assert abs(e - dim * 0.5) <= 0.01 * dim * 0.5
        r, _ = c.evolve(30001, 1, energy_every=2000)
        e2, _ = c.energy()
        assert abs(e2 - e) <= 0.01 * dim * 0.5


def test_ho_ground_state_tight():
    # linear ground state is Gaussian: E = mu = 1 in 2D to O(dt^2)
    with G.Context(dimnum=2, xDim=64, yDim=64, xMax=7, yMax=7, gdt=1e-3) as c:
        c.upload(seeded_psi(c.shape, [c.grid(0), c.grid(1)], seed=4))
        c.parsum()
        c.evolve(4000, 0)
        e, mu = c.energy()
        assert abs(e - 1.0) < 1e-5 and abs(mu - 1.0) < 1e-5
