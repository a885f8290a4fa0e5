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


@pytest.mark.parametrize("prec", [64, 32])
def test_config3_fullsize_gauge_against_oracle(prec):
    # config 3 at its full size: 256^3 over [-8,8)^3, V = (x^2+y^2+4z^2)/2, g = 5000,
    # Omega = 0.6 about z through the gauge operators, dt = 5e-4, from a seeded noisy
    # Gaussian; 4 imaginary steps (both gauge orders twice) against the FP64 oracle.
    # FP64: psi relative L2, E and mu <= 1e-10; FP32 bulk: psi <= 1e-5, E and mu <= 1e-6.
    from oracle import oracle as O
    from tests.conftest import rel_l2

    kw = dict(dimnum=3, xDim=256, yDim=256, zDim=256, xMax=8, yMax=8, zMax=8, g=5000, gpe=1, omegaZ=2.0,
              corotating=1, omega=0.6, gdt=5e-4)
    with G.Context(precision=prec, **kw) as c:
        orc = O.Oracle(**kw)
        psi = seeded_psi(c.shape, [c.grid(a) for a in range(3)], seed=1)
        psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
        c.upload(psi)
        orc.set_wfc(psi)
        del psi
        c.evolve(4, 0)
        orc.evolve(4, 0)
        e_psi = rel_l2(c.download(), orc.wfc())
        (eg, mg), (eo, mo) = c.energy(), orc.energy()
    e_e, e_m = abs(eg - eo) / abs(eo), abs(mg - mo) / abs(mo)
    print(prec, e_psi, e_e, e_m)
    if prec == 64:
        assert e_psi <= 1e-10 and e_e <= 1e-10 and e_m <= 1e-10
    else:
        assert e_psi <= 1e-5 and e_e <= 1e-6 and e_m <= 1e-6


def test_config2_full_job_vortex_lattice_bit_exact():
    # config 2 as the benchmark runs it: 512^2 over [-12,12)^2, g = 1000, Omega = 0.9
    # gauge, 4000 imaginary steps at 1e-3 then 10000 real steps at 1e-4 from the
    # seeded random-phase Gaussian; the vortex markers and list of the final state
    # equal the oracle's findVortex/vortPos on the downloaded psi bit for bit, inside
    # a mask of radius 5 (the Thomas-Fermi radius is ~6)
    from oracle import oracle as O

    kw = dict(dimnum=2, xDim=512, yDim=512, xMax=12, yMax=12, g=1000, gpe=1, corotating=1, omega=0.9, gdt=1e-3,
              dt=1e-4)
    with G.Context(**kw) as c:
        c.set_graph_steps(32)
        psi = seeded_psi(c.shape, [c.grid(0), c.grid(1)], seed=1, phase=True, noise=0.0)
        dV = (c.grid(0)[1] - c.grid(0)[0]) ** 2
        c.upload(psi / np.sqrt(np.sum(np.abs(psi) ** 2) * dV))
        c.evolve(4000, 0)
        c.evolve(10000, 1)
        w = c.download()
        assert abs(np.sum(np.abs(w) ** 2) * dV - 1.0) < 1e-10
        radius = 5.0
        m, found = c.vortex_marker(radius)
        mo, fo = O.find_vortex(w, radius, c.grid(0))
        assert found == fo and np.array_equal(m, mo)
        assert found >= 10   # a lattice has formed
        lst = c.vortex_list(radius)
        assert [tuple(r) for r in lst] == [tuple(r) for r in O.vort_pos(mo)]
