"""Field sources behind the ABI: host tables (GPUE -A file / Grid arrays),
equation strings (static and time dependent, src/dynamic.cu) and the quench."""
import numpy as np
import pytest

import gpue_amd as G
from oracle import eqn_oracle as EO
from oracle import oracle as O
from tests.conftest import rel_l2, seeded_psi

pytestmark = pytest.mark.gpu
TOL64 = 1e-10


def pair(**kw):
    ctx = G.Context(**kw)
    orc = O.Oracle(**{k: v for k, v in kw.items() if k != "a_kind"})
    psi = seeded_psi(ctx.shape, [ctx.grid(a) for a in range(ctx.dimnum)], seed=2, phase=True)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
    ctx.upload(psi)
    orc.set_wfc(psi)
    return ctx, orc


# (48, 1): the strided sweeps' 64-column tiles leave lanes past xDim, which
# must not read the per-point tables out of range
@pytest.mark.parametrize("n,variant", [(64, None), (48, 1)])
def test_table_fields_match_oracle(n, variant):
    kw = dict(dimnum=2, xDim=n, yDim=n, xMax=8, yMax=8, g=40, gpe=1, corotating=1, omega=0.0, gdt=1e-3)
    ctx, orc = pair(**kw)
    if variant is not None:
        ctx.set_variants([variant] * 3)
    x = ctx.grid(0)
    X, Y = np.meshgrid(x, x)
    V = 0.5 * (X ** 2 + 1.3 * Y ** 2) + 0.2 * np.sin(X)
    K = orc.field(O.FIELD_K) * 1.1
    Ax = 0.3 * Y * Y
    Ay = -0.2 * X
    for which, f in [(G.FIELD_V, V), (G.FIELD_K, K), (G.FIELD_AX, Ax), (G.FIELD_AY, Ay)]:
        ctx.set_field(which, f)
        orc.set_field(which, f)
    orc.rebuild()
    ctx.evolve(15, 0)
    orc.evolve(15, 0)
    assert rel_l2(ctx.download(), orc.wfc()) <= TOL64
    ctx.evolve(15, 1)
    orc.evolve(15, 1)
    assert rel_l2(ctx.download(), orc.wfc()) <= TOL64
    (eg, _), (eo, _) = ctx.energy(), orc.energy()
    assert abs(eg - eo) <= TOL64 * abs(eo)


def test_a_table_keeps_harmonic_v_with_half_a_squared():
    # GPUE builds V after A (kharmonic_V adds 1/2 m |A|^2, src/operators.cu:1006-1007)
    n = 32
    kw = dict(dimnum=2, xDim=n, yDim=n, xMax=6, yMax=6, g=20, gpe=1, corotating=1, omega=0.3)
    ctx, orc = pair(**kw)
    x = ctx.grid(0)
    X, Y = np.meshgrid(x, x)
    Ax = 0.1 * Y ** 3
    ctx.set_field(G.FIELD_AX, Ax)
    orc.set_field(O.FIELD_AX, Ax)
    orc.rebuild()
    ctx.evolve(10, 0)
    orc.evolve(10, 0)
    assert rel_l2(ctx.download(), orc.wfc()) <= TOL64


def test_static_equation_field():
    # V from an equation string, evaluated at find_field's coordinates dx*i - xMax
    n, L = 64, 8.0
    kw = dict(dimnum=2, xDim=n, yDim=n, xMax=L, yMax=L, g=30, gpe=1, gdt=1e-3)
    ctx, orc = pair(**kw)
    expr = "0.5*(x*x+y*y)+0.3*cos(x)*sin(y)"
    ctx.set_field_eqn(G.FIELD_V, expr)
    tree = EO.parse(expr)
    dx = 2 * L / n
    V = np.array([[EO.evaluate(tree, dx * i - L, dx * j - L) for i in range(n)] for j in range(n)])
    orc.set_field(O.FIELD_V, V)
    orc.rebuild()
    ctx.evolve(10, 0)
    orc.evolve(10, 0)
    assert rel_l2(ctx.download(), orc.wfc()) <= TOL64


def test_time_dependent_equation_field():
    # a V with t is re-evaluated at time Dt*i before step i (src/evolution.cu:408,732)
    n, L, Dt = 32, 6.0, 1e-2
    kw = dict(dimnum=2, xDim=n, yDim=n, xMax=L, yMax=L, g=10, gpe=1, gdt=Dt, dt=Dt)
    ctx, orc = pair(**kw)
    expr = "0.5*(x*x+y*y)*(1+0.5*sin(3*t))"
    ctx.set_field_eqn(G.FIELD_V, expr)
    tree = EO.parse(expr)
    dx = 2 * L / n
    for gstate in (0, 1):
        for i in range(6):
            t = Dt * i
            V = np.array([[EO.evaluate(tree, dx * a - L, dx * b - L, 0.0, t) for a in range(n)] for b in range(n)])
            orc.set_field(O.FIELD_V, V)
            orc.rebuild()
            orc.step(i, gstate)
        ctx.evolve(6, gstate)
        assert rel_l2(ctx.download(), orc.wfc()) <= TOL64


def test_quench_set_g():
    n = 64
    kw = dict(dimnum=2, xDim=n, yDim=n, xMax=8, yMax=8, g=100, gpe=1, gdt=1e-3, dt=1e-3)
    ctx, orc = pair(**kw)
    ctx.evolve(20, 0)
    orc.evolve(20, 0)
    ctx.set_g(150.0)
    orc.set_g(150.0)
    ctx.evolve(20, 1)
    orc.evolve(20, 1)
    assert rel_l2(ctx.download(), orc.wfc()) <= TOL64


def test_bad_equation_reports_parse_error():
    with G.Context(dimnum=2, xDim=32, yDim=32, xMax=4, yMax=4) as c:
        with pytest.raises(G.GpueError) as ei:
            c.set_field_eqn(G.FIELD_V, "x+")
        assert ei.value.status == G.GPUE_EPARSE
        # a time-dependent program is evaluated with its stack in registers: at most 8 deep
        deep = "x+(x+(x+(x+(x+(x+(x+(x+(x+t))))))))"
        with pytest.raises(G.GpueError) as ei:
            c.set_field_eqn(G.FIELD_AX, deep)
        assert ei.value.status == G.GPUE_EUNSUPPORTED
        c.set_field_eqn(G.FIELD_AX, "x*t")


TD_CASES = {
    # K(t), evaluated like a K table at find_field's coordinates of the grid index
    "K": (dict(dimnum=2, xDim=32, yDim=32, xMax=6, yMax=6), {G.FIELD_K: "0.5*(x*x+y*y)*(1+0.2*sin(5*t))"}),
    # A(t) of a rotating frame whose rate ramps; V = trap + 1/2 m |A(t)|^2 (GPUE builds V after A)
    "A": (dict(dimnum=2, xDim=32, yDim=32, xMax=6, yMax=6),
          {G.FIELD_AX: "0-y*0.3*(1+2*t)", G.FIELD_AY: "x*0.3*(1+2*t)"}),
    "VKA": (dict(dimnum=2, xDim=48, yDim=48, xMax=6, yMax=6),
            {G.FIELD_V: "0.5*(x*x+y*y)*(1+0.5*sin(3*t))", G.FIELD_K: "0.5*(x*x+y*y)*(1+0.1*t)",
             G.FIELD_AX: "0-y*0.3*cos(t)", G.FIELD_AY: "x*0.3"}),
    "3d-VA": (dict(dimnum=3, xDim=16, yDim=16, zDim=16, xMax=5, yMax=5, zMax=5, omegaZ=1.5),
              {G.FIELD_V: "0.5*(x*x+y*y+z*z)+0.2*sin(2*t)*x", G.FIELD_AX: "0-y*0.4*(1+t)",
               G.FIELD_AY: "x*0.4*(1+t)", G.FIELD_AZ: "0.1*x*t"}),
}
ORACLE_FIELD = {G.FIELD_V: O.FIELD_V, G.FIELD_K: O.FIELD_K, G.FIELD_AX: O.FIELD_AX, G.FIELD_AY: O.FIELD_AY,
                G.FIELD_AZ: O.FIELD_AZ}


def _field_at(expr, shape, L, t):
    tree = EO.parse(expr)
    d = [2 * L[a] / shape[::-1][a] for a in range(len(shape))]
    idx = np.indices(shape)   # (z,) y, x
    coords = [d[a] * idx[len(shape) - 1 - a] - L[a] for a in range(len(shape))]
    while len(coords) < 3:
        coords.append(np.zeros(shape))
    f = np.vectorize(lambda x, y, z: EO.evaluate(tree, x, y, z, t))
    return f(*coords)


@pytest.mark.parametrize("case", sorted(TD_CASES))
def test_time_dependent_fields_in_the_sweeps(case):
    # time-dependent V, K and A are evaluated per point inside the step's sweeps
    # at t = Dt*i (no table is written per step); the oracle sets the same fields
    # as tables before each step.  FP64: psi, E and mu <= 1e-10.
    geo, exprs = TD_CASES[case]
    Dt = 1e-2
    kw = dict(geo, g=10, gpe=1, corotating=1, omega=0.3, gdt=Dt, dt=Dt)
    prec = 64
    ctx = G.Context(precision=prec, **kw)
    orc = O.Oracle(**kw)
    psi = seeded_psi(ctx.shape, [ctx.grid(a) for a in range(ctx.dimnum)], seed=3, phase=True)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
    ctx.upload(psi)
    orc.set_wfc(psi)
    for which, e in exprs.items():
        ctx.set_field_eqn(which, e)
    L = [kw["xMax"], kw.get("yMax", 0.0), kw.get("zMax", 0.0)][: ctx.dimnum]
    step = 0
    for gstate, nsteps in ((0, 6), (1, 6)):
        for i in range(step, step + nsteps):
            for which, e in exprs.items():
                orc.set_field(ORACLE_FIELD[which], _field_at(e, ctx.shape, L, Dt * i))
            orc.rebuild()
            orc.step(i, gstate)
        ctx.evolve(nsteps, gstate, first_step=step)
        step += nsteps
        tol = 1e-10 if prec == 64 else 1e-5
        assert rel_l2(ctx.download(), orc.wfc()) <= tol, (case, gstate)
    (eg, mg), (eo, mo) = ctx.energy(), orc.energy()
    tol = 1e-10 if prec == 64 else 1e-6
    assert abs(eg - eo) <= tol * abs(eo) and abs(mg - mo) <= tol * abs(mo), (eg, eo, mg, mo)
    ctx.close()


def test_time_dependent_fields_need_fp64():
    with G.Context(precision=32, dimnum=2, xDim=32, yDim=32, xMax=4, yMax=4) as c:
        with pytest.raises(G.GpueError) as ei:
            c.set_field_eqn(G.FIELD_V, "x*x*t")
        assert ei.value.status == G.GPUE_EUNSUPPORTED
        c.set_field_eqn(G.FIELD_V, "x*x")   # a static field is fine
