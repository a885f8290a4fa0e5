"""Size-independent properties at the benchmark sizes (512^2 and 256^3) and at
the largest state (512^3 FP64, 2 GiB), where the oracle would take minutes per step."""
import numpy as np
import pytest

import gpue_amd as G
from tests.conftest import seeded_psi

pytestmark = pytest.mark.gpu


def test_512_real_time_unitarity_and_energy():
    kw = dict(dimnum=2, xDim=512, yDim=512, xMax=12, yMax=12, g=1000, gpe=1, corotating=1, omega=0.9, gdt=1e-3,
              dt=1e-4)
    with G.Context(**kw) as c:
        c.upload(seeded_psi(c.shape, [c.grid(0), c.grid(1)], seed=1, phase=True, noise=0.0))
        c.parsum()
        c.evolve(300, 0)
        e0, _ = c.energy()
        c.evolve(2000, 1)
        w = c.download()
        dV = (c.grid(0)[1] - c.grid(0)[0]) ** 2
        assert abs(np.sum(np.abs(w) ** 2) * dV - 1.0) < 1e-11
        e1, _ = c.energy()
        assert abs(e1 - e0) < 1e-3 * abs(e0)


def test_256cube_fp32_vs_fp64_and_monotone_energy():
    kw = dict(dimnum=3, xDim=256, yDim=256, zDim=256, xMax=8, yMax=8, zMax=8, g=5000, gpe=1, omegaZ=2.0, gdt=5e-4)
    res = {}
    for prec in (64, 32):
        with G.Context(precision=prec, **kw) as c:
            psi = seeded_psi(c.shape, [c.grid(a) for a in range(3)], seed=1)
            c.upload(psi)
            c.parsum()
            r, en = c.evolve(60, 0, energy_every=20)
            assert np.all(np.diff(en) < 0)          # imaginary time lowers the energy
            res[prec] = c.energy()
    for k in range(2):
        assert abs(res[32][k] - res[64][k]) <= 1e-6 * abs(res[64][k])


def test_512cube_largest_state_properties():
    # 512^3 FP64: psi is exactly 2^31 bytes (every byte offset past 2 GiB needs 64-bit
    # arithmetic) and the reducing sweeps leave ~10^5 partials.  Imaginary time lowers the
    # energy; real time keeps the norm and conserves the energy of the evolved
    # Hamiltonian (the oracle on 64^3 with the same operators: 1.4e-11 over 20 steps);
    # FP32 bulk agrees with FP64 to 1e-6.
    kw = dict(dimnum=3, xDim=512, yDim=512, zDim=512, xMax=8, yMax=8, zMax=8, g=2000, gpe=1, omegaZ=1.5,
              corotating=1, omega=0.5, gdt=1e-3, dt=1e-4)
    res = {}
    for prec in (64, 32):
        with G.Context(precision=prec, **kw) as c:
            x, y, z = c.grid(0), c.grid(1), c.grid(2)
            gzy = np.exp(-(z ** 2) / 2)[:, None] * np.exp(-((y - 0.4) ** 2) / 2)[None, :]
            gx = np.exp(-(x ** 2) / 2) * np.exp(0.3j * x)
            c.upload(gzy[:, :, None] * gx[None, None, :])
            c.parsum()
            r, en = c.evolve(40, 0, energy_every=10)
            assert r.steps_done == 40 and len(en) == 4
            assert np.all(np.diff(en) < 0)
            res[prec] = c.energy()
            if prec == 64:
                e0, _ = c.energy_consistent()
                c.evolve(20, 1)
                w = c.download()
                dV = (x[1] - x[0]) ** 3
                assert abs(np.sum(np.abs(w) ** 2) * dV - 1.0) < 1e-11
                del w
                e1, _ = c.energy_consistent()
                assert abs(e1 - e0) < 1e-8 * abs(e0)
    for k in range(2):
        assert abs(res[32][k] - res[64][k]) <= 1e-6 * abs(res[64][k])
