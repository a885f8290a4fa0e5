"""Energy with the gauge term of the evolved Hamiltonian (gpue_energy_consistent)
against the FP64 oracle (or_energy_consistent).

GPUE's energy_calc adds +i hbar A.grad by forward differences (src/split_op.cu:312-360);
the step evolves exp(-A_a k_a Dt) between 1-D transforms (src/operators.cu:1089-1110),
i.e. +hbar A.k.  gpue_energy_consistent reports the latter, spectrally.

Tolerances: FP64 energy and mu relative <= 1e-10; FP32 bulk <= 1e-6 (as gpue_energy).
"""
import numpy as np
import pytest

import gpue_amd as G
from oracle import oracle as O
from tests.conftest import seeded_psi

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-6


def _pair(prec=64, steps=6, tables=None, **kw):
    ctx = G.Context(precision=prec, **kw)
    orc = O.Oracle(**kw)
    if tables:
        for which, f in tables(ctx):
            ctx.set_field(which, f)
            orc.set_field(which, f)
        orc.rebuild()
    grids = [ctx.grid(a) for a in range(ctx.dimnum)]
    psi = seeded_psi(ctx.shape, grids, seed=5, phase=True)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
    ctx.upload(psi)
    orc.set_wfc(psi)
    if steps:
        ctx.evolve(steps, 0)
        for i in range(steps):
            orc.step(i, 0)
    return ctx, orc


def _xdep_a(ctx):
    # Ax depends on x and y, Ay on x and y: the mixed-representation convention of
    # the step (A_a[gid] with the k index along a) decides the value
    X, Y = np.meshgrid(ctx.grid(0), ctx.grid(1))
    return [(G.FIELD_AX, -0.4 * Y + 0.05 * X * Y), (G.FIELD_AY, 0.5 * X + 0.03 * X * X)]


def _az_table(ctx):
    z, y, x = np.meshgrid(ctx.grid(2), ctx.grid(1), ctx.grid(0), indexing="ij")
    return [(G.FIELD_AX, -0.3 * y), (G.FIELD_AY, 0.3 * x), (G.FIELD_AZ, 0.2 * x + 0.1 * y)]


CASES = [
    ("2d-gauge", dict(dimnum=2, xDim=64, yDim=64, xMax=8, yMax=8, g=50, gpe=1, corotating=1, omega=0.5), None),
    ("2d-aniso", dict(dimnum=2, xDim=128, yDim=128, xMax=9, yMax=9, g=80, gpe=1, corotating=1, omega=0.4,
                      omegaX=1.0, omegaY=1.4, gammaY=1.15), None),
    # rectangular and Az-table cases check the energy alone (no steps before it)
    ("2d-rect", dict(dimnum=2, xDim=128, yDim=64, xMax=10, yMax=5, g=50, gpe=1, corotating=1, omega=0.3,
                     steps=0), None),
    ("2d-hbar", dict(dimnum=2, xDim=64, yDim=64, xMax=8, yMax=8, g=30, gpe=1, corotating=1, omega=0.6,
                     hbar=1.7, mass=0.8), None),
    ("2d-atable", dict(dimnum=2, xDim=64, yDim=64, xMax=8, yMax=8, g=50, gpe=1, corotating=1, omega=0.5), _xdep_a),
    ("3d-gauge", dict(dimnum=3, xDim=32, yDim=32, zDim=32, xMax=6, yMax=6, zMax=6, g=100, gpe=1, omegaZ=2.0,
                      corotating=1, omega=0.6), None),
    ("3d-azt", dict(dimnum=3, xDim=32, yDim=48, zDim=16, xMax=6, yMax=6, zMax=4, g=60, gpe=1, omegaZ=1.5,
                    corotating=1, omega=0.3, steps=0), _az_table),
]


@pytest.mark.parametrize("name,kw,tables", CASES, ids=[c[0] for c in CASES])
def test_energy_consistent_fp64(name, kw, tables):
    kw = dict(kw)
    ctx, orc = _pair(tables=tables, steps=kw.pop("steps", 6), **kw)
    eg, mg = ctx.energy_consistent()
    eo, mo = orc.energy_consistent()
    print(name, eg, eo, mg, mo)
    assert abs(eg - eo) <= TOL64 * abs(eo)
    assert abs(mg - mo) <= TOL64 * abs(mo)
    # and GPUE's own energy is still the forward-difference one
    e_gpue, _ = ctx.energy()
    assert abs(e_gpue - orc.energy()[0]) <= TOL64 * abs(eo)
    ctx.close()


def test_energy_consistent_fp32():
    kw = dict(dimnum=2, xDim=128, yDim=128, xMax=9, yMax=9, g=80, gpe=1, corotating=1, omega=0.4)
    ctx, orc = _pair(prec=32, **kw)
    eg, mg = ctx.energy_consistent()
    eo, mo = orc.energy_consistent()
    assert abs(eg - eo) <= TOL32 * abs(eo)
    assert abs(mg - mo) <= TOL32 * abs(mo)
    ctx.close()


@pytest.mark.parametrize("kw", [
    dict(dimnum=2, xDim=64, yDim=64, xMax=8, yMax=8, g=50, gpe=1),
    dict(dimnum=1, xDim=128, xMax=8, g=10, gpe=1, corotating=1, omega=0.3),
], ids=["2d-no-gauge", "1d"])
def test_energy_consistent_equals_energy_without_gauge_term(kw):
    ctx, _ = _pair(**kw)
    assert ctx.energy_consistent() == ctx.energy()
    ctx.close()
