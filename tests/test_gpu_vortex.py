"""Vortex plaquette detection (Tracker::findVortex, src/tracker.cu:102-161):
integer output, bit-exact against the geometry fixtures and the oracle."""
import numpy as np
import pytest

import gpue_amd as G
from oracle import oracle as O
from tests.conftest import seeded_psi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [64, 32])
def test_vortex_fixtures(golden, prec):
    d = golden("vortex_cases.npz")
    n = d["v1_psi"].shape[0]
    with G.Context(dimnum=2, xDim=n, yDim=n, xMax=8.0, yMax=8.0, precision=prec) as c:
        assert np.array_equal(c.grid(0), d["x"])
        for case in ["v1", "v2", "v3"]:
            c.upload(d[case + "_psi"])
            m, found = c.vortex_marker(float(d[case + "_radius"]))
            assert np.array_equal(m, d[case + "_marker"]), case
            assert found == int(np.count_nonzero(d[case + "_marker"])), case
        c.upload(d["v1_psi"])
        lst = c.vortex_list(100.0)
        assert [tuple(r) for r in lst] == [(20, 20, 1), (20, 40, 1), (40, 30, 1), (45, 45, 1)]


def test_vortex_random_phase_matches_oracle():
    # a random phase field has candidates almost everywhere, which exercises the
    # sequential skip rule on long runs of consecutive candidates
    n = 128
    rng = np.random.default_rng(3)
    with G.Context(dimnum=2, xDim=n, yDim=n, xMax=10.0, yMax=10.0) as c:
        x = c.grid(0)
        for trial in range(3):
            psi = np.exp(1j * rng.uniform(0, 2 * np.pi, (n, n))) * (1 + 0.1 * rng.uniform(size=(n, n)))
            c.upload(psi)
            for radius in (3.0, 7.5, 100.0):
                m, found = c.vortex_marker(radius)
                mo, fo = O.find_vortex(psi, radius, x)
                assert found == fo
                assert np.array_equal(m, mo)


def test_vortex_lattice_after_evolution():
    # rotating condensate (config-2 physics on 256^2): evolve on the GPU, then the GPU
    # detector and the oracle detector see the same psi and must agree bit for bit
    kw = dict(dimnum=2, xDim=256, yDim=256, xMax=12, yMax=12, g=1000, gpe=1, corotating=1, omega=0.9, gdt=1e-3,
              dt=1e-4)
    with G.Context(**kw) as c:
        psi = seeded_psi(c.shape, [c.grid(0), c.grid(1)], seed=1, phase=True, noise=0.0)
        c.upload(psi)
        c.parsum()
        c.evolve(1500, 0)
        c.evolve(200, 1)
        w = c.download()
        for radius in (4.0, 6.0):
            m, found = c.vortex_marker(radius)
            mo, fo = O.find_vortex(w, radius, c.grid(0))
            assert found == fo and found > 0
            assert np.array_equal(m, mo)
            lst = c.vortex_list(radius)
            assert [tuple(r) for r in lst] == [tuple(r) for r in O.vort_pos(mo)]


def test_vortex_rejects_non_square():
    with G.Context(dimnum=2, xDim=64, yDim=32, xMax=8, yMax=4) as c:
        with pytest.raises(G.GpueError):
            c.vortex_marker(2.0)
    with G.Context(dimnum=3, xDim=16, yDim=16, zDim=16, xMax=4, yMax=4, zMax=4) as c:
        with pytest.raises(G.GpueError):
            c.vortex_marker(2.0)
