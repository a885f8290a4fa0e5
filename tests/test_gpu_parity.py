"""HIP path vs the FP64 oracle and the golden fixtures.

Tolerances (stated here, checked below):
  FP64:  psi relative L2 <= 1e-10; energy and chemical potential relative <= 1e-10
  FP32 bulk / FP64 reductions: energy and mu relative <= 1e-6; psi relative L2 <= 1e-5,
    in imaginary and in real time (the real-time FP32 kernels form the static V, A and
    K phase factors and their products in FP64, so no FP32-rounded factor repeats
    its modulus error step after step)
"""
import numpy as np
import pytest

import gpue_amd as G
from oracle import oracle as O
from tests.conftest import rel_l2, seeded_psi

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32_E = 1e-6
TOL32_PSI = 1e-5


def run_pair(nsteps, gstate, prec=64, phase=True, seed=1, first_step=0, **kw):
    ctx = G.Context(precision=prec, **kw)
    orc = O.Oracle(**{k: v for k, v in kw.items() if k != "a_kind"})
    grids = [ctx.grid(a) for a in range(ctx.dimnum)]
    psi = seeded_psi(ctx.shape, grids, seed=seed, phase=phase)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
    ctx.upload(psi)
    orc.set_wfc(psi)
    r, _ = ctx.evolve(nsteps, gstate, first_step=first_step)
    assert r.steps_done == nsteps
    for i in range(first_step, first_step + nsteps):
        orc.step(i, gstate)
    a, b = ctx.download(), orc.wfc()
    eg, mg = ctx.energy()
    eo, mo = orc.energy()
    ctx.close()
    return rel_l2(a, b), abs(eg - eo) / abs(eo), abs(mg - mo) / abs(mo)


CASES = [
    ("1d", 1, dict(dimnum=1, xDim=64, xMax=8, g=10, gpe=1)),
    ("1d-gauge", 1, dict(dimnum=1, xDim=128, xMax=8, g=10, gpe=1, corotating=1, omega=0.3)),
    ("2d-plain", 2, dict(dimnum=2, xDim=64, yDim=64, xMax=8, yMax=8, g=50, gpe=1)),
    ("2d-linear", 2, dict(dimnum=2, xDim=32, yDim=32, xMax=6, yMax=6)),
    ("2d-gauge", 2, dict(dimnum=2, xDim=64, yDim=64, xMax=8, yMax=8, g=50, gpe=1, corotating=1, omega=0.5)),
    ("2d-aniso", 2, dict(dimnum=2, xDim=128, yDim=128, xMax=9, yMax=9, g=80, gpe=1, corotating=1, omega=0.4,
                         gammaY=1.15, x0_shift=0.3, y0_shift=-0.2)),
    ("2d-192", 2, dict(dimnum=2, xDim=192, yDim=192, xMax=10, yMax=10, g=200, gpe=1, corotating=1, omega=0.7)),
    ("2d-384", 2, dict(dimnum=2, xDim=384, yDim=384, xMax=10, yMax=10, g=200, gpe=1, corotating=1, omega=0.7)),
    ("2d-1024", 2, dict(dimnum=2, xDim=1024, yDim=1024, xMax=12, yMax=12, g=400, gpe=1, corotating=1, omega=0.7)),
    ("2d-rect", 2, dict(dimnum=2, xDim=128, yDim=64, xMax=10, yMax=5, g=50, gpe=1)),
    ("3d-plain", 3, dict(dimnum=3, xDim=32, yDim=32, zDim=32, xMax=6, yMax=6, zMax=6, g=100, gpe=1, omegaZ=2.0)),
    ("3d-gauge", 3, dict(dimnum=3, xDim=32, yDim=32, zDim=32, xMax=6, yMax=6, zMax=6, g=100, gpe=1, omegaZ=2.0,
                         corotating=1, omega=0.6)),
    ("3d-box", 3, dict(dimnum=3, xDim=64, yDim=32, zDim=16, xMax=8, yMax=6, zMax=4, g=60, gpe=1, omegaZ=1.5)),
    ("2d-48", 2, dict(dimnum=2, xDim=48, yDim=48, xMax=6, yMax=6, g=50, gpe=1, corotating=1, omega=0.5)),
    ("3d-96", 3, dict(dimnum=3, xDim=96, yDim=48, zDim=48, xMax=8, yMax=6, zMax=6, g=60, gpe=1, omegaZ=1.5,
                      corotating=1, omega=0.4)),
    # long strided lines along z: 1024 (radix 16,16,4 / 8,8,8,2) and 384 (12,8,4 / 12,4,4,2)
    ("3d-longz", 3, dict(dimnum=3, xDim=32, yDim=16, zDim=1024, xMax=5, yMax=4, zMax=40, g=40, gpe=1, omegaZ=0.05,
                         corotating=1, omega=0.3)),
    ("3d-384z", 3, dict(dimnum=3, xDim=32, yDim=48, zDim=384, xMax=6, yMax=6, zMax=20, g=60, gpe=1, omegaZ=0.3,
                        corotating=1, omega=0.4)),
    # the largest lines: 2048 (radix 16,16,8 / 8,8,8,4) and 4096 (16,16,16 / 8,8,8,8), along x (rows),
    # y (strided, stride xDim) and z (strided, stride xDim*yDim)
    ("2d-2048", 2, dict(dimnum=2, xDim=2048, yDim=2048, xMax=14, yMax=14, g=600, gpe=1, corotating=1, omega=0.7)),
    ("2d-4096x64", 2, dict(dimnum=2, xDim=4096, yDim=64, xMax=40, yMax=6, g=100, gpe=1, corotating=1, omega=0.3)),
    ("2d-64x4096", 2, dict(dimnum=2, xDim=64, yDim=4096, xMax=6, yMax=40, g=100, gpe=1, corotating=1, omega=0.3,
                           gammaY=0.05)),
    ("3d-2048z", 3, dict(dimnum=3, xDim=16, yDim=16, zDim=2048, xMax=4, yMax=4, zMax=60, g=40, gpe=1, omegaZ=0.02,
                         corotating=1, omega=0.3)),
    ("3d-4096z", 3, dict(dimnum=3, xDim=16, yDim=16, zDim=4096, xMax=4, yMax=4, zMax=80, g=40, gpe=1, omegaZ=0.01,
                         corotating=1, omega=0.3)),
]


@pytest.mark.parametrize("gstate", [0, 1])
@pytest.mark.parametrize("name,dim,kw", CASES, ids=[c[0] for c in CASES])
def test_fp64_parity(name, dim, kw, gstate):
    steps = 4 if kw.get("xDim", 0) >= 1024 else 12
    e_psi, e_e, e_mu = run_pair(steps, gstate, first_step=3, gdt=1e-3, dt=5e-4, **kw)
    assert e_psi <= TOL64
    assert e_e <= TOL64 and e_mu <= TOL64


FP32_CASES = [c for c in CASES if c[0] in ("2d-gauge", "3d-gauge", "2d-192", "1d")]


@pytest.mark.parametrize("name,dim,kw", FP32_CASES, ids=[c[0] for c in FP32_CASES])
@pytest.mark.parametrize("gstate", [0, 1])
def test_fp32_bulk_parity(name, dim, kw, gstate):
    e_psi, e_e, e_mu = run_pair(20, gstate, prec=32, gdt=1e-3, dt=5e-4, **kw)
    print(name, gstate, e_psi, e_e, e_mu)
    assert e_psi <= TOL32_PSI
    assert e_e <= TOL32_E and e_mu <= TOL32_E


def test_matlab_crosscheck_fixture(golden):
    g = golden("gpe2d_matlab.npz")
    n = int(g["n"])
    ctx = G.Context(dimnum=2, xDim=n, yDim=n, xMax=float(g["xmax"]), yMax=float(g["xmax"]), g=float(g["g"]), gpe=1,
                    gdt=float(g["Dt"]))
    ctx.upload(g["psi0"])
    ctx.evolve(1, 0)
    assert rel_l2(ctx.download(), g["psi_1"]) <= TOL64
    ctx.evolve(int(g["steps"]) - 1, 0, first_step=1)
    assert rel_l2(ctx.download(), g["psi_%d" % int(g["steps"])]) <= TOL64


@pytest.mark.parametrize("dim", [2, 3])
def test_gauge_order_fixture(golden, dim):
    g = golden(f"gpue_gauge_{dim}d.npz")
    n, L = int(g["n"]), float(g["xmax"])
    ctx = G.Context(dimnum=dim, xDim=n, yDim=n, zDim=n, xMax=L, yMax=L, zMax=L, g=float(g["g"]), gpe=1, corotating=1,
                    omega=float(g["omega"]), omegaZ=2.0, gdt=float(g["gdt"]), dt=float(g["dt"]))
    ctx.upload(g["psi0"])
    ctx.evolve(int(g["steps"]), 0)
    assert rel_l2(ctx.download(), g["psi_imag"]) <= TOL64
    ctx.evolve(int(g["steps"]), 1)
    assert rel_l2(ctx.download(), g["psi_real"]) <= TOL64


def test_config1_ground_state_full():
    # 256^2 over [-10,10)^2, V = (x^2+y^2)/2, g = 500, dt = 1e-3, 2000 imaginary steps,
    # psi0 = exp(-r^2/2)(1 + 0.01 U[-1,1]) seed 1, normalised
    kw = dict(dimnum=2, xDim=256, yDim=256, xMax=10, yMax=10, g=500, gpe=1, gdt=1e-3)
    ctx = G.Context(**kw)
    orc = O.Oracle(**kw)
    grids = [ctx.grid(0), ctx.grid(1)]
    psi = seeded_psi(ctx.shape, grids, seed=1)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
    ctx.upload(psi)
    orc.set_wfc(psi)
    ctx.evolve(2000, 0)
    orc.evolve(2000, 0)
    assert rel_l2(ctx.download(), orc.wfc()) <= TOL64
    (eg, mg), (eo, mo) = ctx.energy(), orc.energy()
    assert abs(eg - eo) <= TOL64 * abs(eo) and abs(mg - mo) <= TOL64 * abs(mo)


def test_config2_rotating_reduced():
    # config 2 at full grid (512^2, [-12,12)^2, g = 1000, Omega = 0.9 gauge), shortened to
    # 100 imaginary steps at 1e-3 then 100 real steps at 1e-4 from a random-phase Gaussian
    kw = dict(dimnum=2, xDim=512, yDim=512, xMax=12, yMax=12, g=1000, gpe=1, corotating=1, omega=0.9, gdt=1e-3,
              dt=1e-4)
    ctx = G.Context(**kw)
    orc = O.Oracle(**kw)
    grids = [ctx.grid(0), ctx.grid(1)]
    psi = seeded_psi(ctx.shape, grids, seed=1, phase=True, noise=0.0)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
    ctx.upload(psi)
    orc.set_wfc(psi)
    ctx.evolve(100, 0)
    orc.evolve(100, 0)
    assert rel_l2(ctx.download(), orc.wfc()) <= TOL64
    ctx.evolve(100, 1)
    orc.evolve(100, 1)
    assert rel_l2(ctx.download(), orc.wfc()) <= TOL64
    (eg, mg), (eo, mo) = ctx.energy(), orc.energy()
    assert abs(eg - eo) <= TOL64 * abs(eo) and abs(mg - mo) <= TOL64 * abs(mo)


@pytest.mark.parametrize("prec,tol", [(64, TOL64), (32, TOL32_E)])
def test_config3_reduced(prec, tol):
    # config 3 physics (V = (x^2+y^2+4z^2)/2, g = 5000, Omega = 0.6 gauge, dt = 5e-4)
    # on 64^3 for 30 imaginary steps
    kw = dict(dimnum=3, xDim=64, yDim=64, zDim=64, xMax=8, yMax=8, zMax=8, g=5000, gpe=1, omegaZ=2.0, corotating=1,
              omega=0.6, gdt=5e-4)
    ctx = G.Context(precision=prec, **kw)
    orc = O.Oracle(**kw)
    grids = [ctx.grid(a) for a in range(3)]
    psi = seeded_psi(ctx.shape, grids, seed=1)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
    ctx.upload(psi)
    orc.set_wfc(psi)
    ctx.evolve(30, 0)
    orc.evolve(30, 0)
    (eg, mg), (eo, mo) = ctx.energy(), orc.energy()
    assert abs(eg - eo) <= tol * abs(eo) and abs(mg - mo) <= tol * abs(mo)
    assert rel_l2(ctx.download(), orc.wfc()) <= (TOL64 if prec == 64 else TOL32_PSI)


def test_energy_cadence_and_early_stop():
    kw = dict(dimnum=2, xDim=64, yDim=64, xMax=6, yMax=6, g=20, gpe=1, gdt=1e-2)
    ctx = G.Context(**kw)
    orc = O.Oracle(**kw)
    grids = [ctx.grid(0), ctx.grid(1)]
    psi = seeded_psi(ctx.shape, grids, seed=9)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
    ctx.upload(psi)
    orc.set_wfc(psi)
    r, eg = ctx.evolve(3000, 0, energy_every=10, thr=1e-7)
    done, eo = orc.evolve(3000, 0, energy_every=10, thr=1e-7)
    assert r.early_stop == 1 and r.steps_done == done
    assert len(eg) == len(eo)
    assert np.max(np.abs(eg - eo) / np.abs(eo)) <= TOL64


@pytest.mark.parametrize("graph", [0, 8])
def test_real_time_after_imaginary_time(graph):
    # imaginary time leaves its last renormalisation pending on the device;
    # real time must start from the renormalised state and never divide again.
    # psi0 is far from normalised so a missed or repeated division shows.
    kw = dict(dimnum=2, xDim=64, yDim=64, xMax=8, yMax=8, g=80, gpe=1, corotating=1, omega=0.5, gdt=1e-3, dt=5e-4)
    ctx = G.Context(**kw)
    ctx.set_graph_steps(graph)
    orc = O.Oracle(**kw)
    psi = 3.0 * seeded_psi(ctx.shape, [ctx.grid(0), ctx.grid(1)], seed=9, phase=True)
    ctx.upload(psi)
    orc.set_wfc(psi)
    ctx.evolve(5, 0)
    ctx.evolve(24, 1, first_step=5)
    for i in range(5):
        orc.step(i, 0)
    for i in range(5, 29):
        orc.step(i, 1)
    assert rel_l2(ctx.download(), orc.wfc()) <= TOL64
    ctx.close()


GRAPH_CASES = {
    "2d-gauge": dict(dimnum=2, xDim=128, yDim=128, xMax=10, yMax=10, g=200, gpe=1, corotating=1, omega=0.7),
    "2d-plain": dict(dimnum=2, xDim=128, yDim=128, xMax=10, yMax=10, g=200, gpe=1),
    "3d-plain": dict(dimnum=3, xDim=32, yDim=32, zDim=32, xMax=6, yMax=6, zMax=6, g=100, gpe=1, omegaZ=2.0),
    "3d-gauge": dict(dimnum=3, xDim=32, yDim=32, zDim=32, xMax=6, yMax=6, zMax=6, g=100, gpe=1, omegaZ=2.0,
                     corotating=1, omega=0.6),
}


@pytest.mark.parametrize("case", sorted(GRAPH_CASES))
@pytest.mark.parametrize("gstate", [0, 1])
def test_graph_replay_matches_eager(gstate, case):
    # steady-state steps replayed from a captured HIP graph must give the
    # eagerly launched result; replays are interleaved with eager work and
    # syncs.  Imaginary time: bit-exact.  Real time: the graph fuses the row
    # sweeps at step boundaries into one kernel, whose code the compiler may
    # contract into FMAs differently from the two separate kernels, so the bar
    # is 1e-13 relative (measured: ~1e-15 after 8 steps, both paths ~1e-15
    # from the oracle).
    kw = dict(GRAPH_CASES[case], gdt=1e-3, dt=5e-4)
    a = G.Context(**kw)
    b = G.Context(**kw)
    a.set_graph_steps(8)
    b.set_graph_steps(0)
    psi = seeded_psi(a.shape, [a.grid(ax) for ax in range(a.dimnum)], seed=5, phase=True)
    for c in (a, b):
        c.upload(psi)
        c.parsum()
    step = 0
    for rep in range(5):
        ra, _ = a.evolve(17, gstate, first_step=step)
        rb, _ = b.evolve(17, gstate, first_step=step)
        step += 17
        wa, wb = a.download(), b.download()
        if gstate == 0:
            assert np.array_equal(wa, wb), rep
            assert a.energy() == b.energy()
        else:
            assert rel_l2(wa, wb) <= 1e-13, rep
            (ea, ma), (eb, mb) = a.energy(), b.energy()
            assert abs(ea - eb) <= 1e-13 * abs(eb) and abs(ma - mb) <= 1e-13 * abs(mb)
    a.close()
    b.close()


VARIANT_CASES = ("1d-gauge", "2d-gauge", "3d-gauge", "2d-48", "3d-96", "3d-longz", "2d-4096x64", "2d-64x4096",
                 "3d-2048z")


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("name,dim,kw", [c for c in CASES if c[0] in VARIANT_CASES],
                         ids=[c[0] for c in CASES if c[0] in VARIANT_CASES])
def test_every_kernel_variant(name, dim, kw, variant):
    # the autotuner may pick any plan and strided tile width per axis: each
    # must meet the FP64 bar, imaginary and real time
    for gstate in (0, 1):
        ctx = G.Context(precision=64, gdt=1e-3, dt=5e-4, **kw)
        ctx.set_variants([variant] * 3)
        orc = O.Oracle(gdt=1e-3, dt=5e-4, **{k: v for k, v in kw.items() if k != "a_kind"})
        psi = seeded_psi(ctx.shape, [ctx.grid(a) for a in range(ctx.dimnum)], seed=2, phase=True)
        psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
        ctx.upload(psi)
        orc.set_wfc(psi)
        ctx.evolve(10, gstate, first_step=1)
        for i in range(1, 11):
            orc.step(i, gstate)
        assert rel_l2(ctx.download(), orc.wfc()) <= TOL64, (variant, gstate)
        ctx.close()


@pytest.mark.parametrize("ramp_type", [1, 2])
@pytest.mark.parametrize("name,dim,kw", [c for c in CASES if c[0] in ("2d-gauge", "3d-gauge")],
                         ids=["2d-gauge", "3d-gauge"])
def test_ramp_matches_oracle(name, dim, kw, ramp_type):
    # -R: the oracle raises GPUE's gauge operators to omega_0 step by step
    # (imaginary time) or scales the phase (real time); the HIP path scales the
    # exponent by the closed-form product.  FP64 bar.
    for gstate in (0, 1):
        n = 12
        ctx = G.Context(precision=64, gdt=1e-3, dt=5e-4, **kw)
        orc = O.Oracle(gdt=1e-3, dt=5e-4, **kw)
        ctx.set_ramp(ramp_type, n)
        orc.set_ramp(ramp_type, n)
        psi = seeded_psi(ctx.shape, [ctx.grid(a) for a in range(ctx.dimnum)], seed=4, phase=True)
        psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
        ctx.upload(psi)
        orc.set_wfc(psi)
        ctx.evolve(n, gstate)
        for i in range(n):
            orc.step(i, gstate)
        assert rel_l2(ctx.download(), orc.wfc()) <= TOL64, (ramp_type, gstate)
        # the ramp changes the result (the test is not vacuous)
        ref = O.Oracle(gdt=1e-3, dt=5e-4, **kw)
        ref.set_wfc(psi)
        for i in range(n):
            ref.step(i, gstate)
        assert rel_l2(ref.wfc(), orc.wfc()) > 1e-6
        ctx.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("name,dim,kw", [c for c in CASES if c[0] in ("2d-gauge", "3d-gauge")],
                         ids=["2d-gauge", "3d-gauge"])
def test_every_kernel_variant_fp32(name, dim, kw, variant):
    # the FP32 sweeps of every variant (split twiddles, 45-degree rotations,
    # FP32 tables) meet the FP32 bar in imaginary and real time
    for gstate in (0, 1):
        ctx = G.Context(precision=32, gdt=1e-3, dt=5e-4, **kw)
        ctx.set_variants([variant] * 3)
        orc = O.Oracle(gdt=1e-3, dt=5e-4, **{k: v for k, v in kw.items() if k != "a_kind"})
        psi = seeded_psi(ctx.shape, [ctx.grid(a) for a in range(ctx.dimnum)], seed=4, phase=True)
        psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
        ctx.upload(psi)
        orc.set_wfc(psi)
        ctx.evolve(20, gstate)
        for i in range(20):
            orc.step(i, gstate)
        assert rel_l2(ctx.download(), orc.wfc()) <= TOL32_PSI, (variant, gstate)
        (eg, mg), (eo, mo) = ctx.energy(), orc.energy()
        assert abs(eg - eo) <= TOL32_E * abs(eo) and abs(mg - mo) <= TOL32_E * abs(mo), (variant, gstate)
        ctx.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_autotune_leaves_state(prec):
    # the autotuner times sweeps and whole steps on copies: after it, the run
    # continues exactly as one that was given the chosen variants directly
    kw = dict(dimnum=2, xDim=128, yDim=128, xMax=9, yMax=9, g=80, gpe=1, corotating=1, omega=0.4,
              gdt=1e-3, dt=5e-4)
    a, b = G.Context(precision=prec, **kw), G.Context(precision=prec, **kw)
    psi = seeded_psi(a.shape, [a.grid(k) for k in range(2)], seed=6, phase=True)
    for c in (a, b):
        c.upload(psi)
        c.evolve(10, 0)   # leaves a norm pending
    v = a.autotune()
    b.set_variants(list(v))
    a.set_variants(list(v))
    wa, wb = a.download(), b.download()   # both settle the pending norm the same way
    assert np.array_equal(wa, wb)
    for c in (a, b):
        c.evolve(16, 0, first_step=10)
        c.evolve(16, 1, first_step=26)
    assert np.array_equal(a.download(), b.download())
    assert a.energy() == b.energy()
    a.close()
    b.close()


def test_fp32_real_time_long_run_does_not_drift():
    # FP32-bulk real time over 2000 steps (graphs on) against the FP64 path: the
    # relative energy and mu errors stay within the FP32 bar, and the error after
    # 2000 steps is far below 100x the error after 20 steps (a phase factor whose
    # modulus error repeats every step would make it grow linearly)
    kw = dict(dimnum=2, xDim=128, yDim=128, xMax=10, yMax=10, g=200, gpe=1, corotating=1, omega=0.7, gdt=1e-3,
              dt=5e-4)
    a, b = G.Context(precision=64, **kw), G.Context(precision=32, **kw)
    psi = seeded_psi(a.shape, [a.grid(0), a.grid(1)], seed=3)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * (a.grid(0)[1] - a.grid(0)[0]) ** 2)
    errs = {}
    for c in (a, b):
        c.set_graph_steps(32)
        c.upload(psi)
    done = 0
    for target in (20, 2000):
        for c in (a, b):
            c.evolve(target - done, 1, first_step=done)
        done = target
        (ea, ma), (eb, mb) = a.energy(), b.energy()
        errs[target] = (abs(eb - ea) / abs(ea), abs(mb - ma) / abs(ma))
        print(target, errs[target])
        assert errs[target][0] <= TOL32_E and errs[target][1] <= TOL32_E, (target, errs[target])
    floor = 1e-8
    assert errs[2000][0] <= 30 * max(errs[20][0], floor), errs
    a.close()
    b.close()


@pytest.mark.parametrize("change", ["set_g", "set_variants", "set_graph_steps"])
def test_parameter_change_right_after_graph_replays(change):
    # a real-time evolve returns with graph replays still queued; changing a
    # parameter destroys the captured graphs, which must wait for them.  The
    # result equals a run that synchronised before the change.
    kw = dict(dimnum=2, xDim=64, yDim=64, xMax=8, yMax=8, g=80, gpe=1, corotating=1, omega=0.5, gdt=1e-3, dt=5e-4)
    psi = None
    out = []
    for sync_first in (False, True):
        c = G.Context(**kw)
        c.set_graph_steps(8)
        if psi is None:
            psi = seeded_psi(c.shape, [c.grid(0), c.grid(1)], seed=7, phase=True)
        c.upload(psi)
        c.evolve(64, 1)
        if sync_first:
            c.sync()
        if change == "set_g":
            c.set_g(90.0)
        elif change == "set_variants":
            c.set_variants([1, 1, 1])
        else:
            c.set_graph_steps(4)
        c.evolve(16, 1, first_step=64)
        out.append(c.download())
        c.close()
    assert rel_l2(out[0], out[1]) <= 1e-13


@pytest.mark.parametrize("case", ["line", "rect", "box"])
def test_gauge_box_fixture(golden, case):
    # the HIP path against the numpy transcription of GPUE's gauge step on a 1D line and on
    # boxes whose axes differ in length and extent (tests/golden/make_golden.py)
    g = golden("gpue_gauge_box.npz")
    dims = [int(v) for v in g[f"{case}_dims"]]
    maxes = [float(v) for v in g[f"{case}_maxes"]]
    gg, om, gam, oz, steps, gdt, dt = [float(v) for v in g[f"{case}_par"]]
    kw = dict(dimnum=len(dims), g=gg, gpe=1, corotating=1, omega=om, gammaY=gam, omegaZ=oz, gdt=gdt, dt=dt)
    for a, (n, m) in enumerate(zip(dims, maxes)):
        kw["xyz"[a] + "Dim"] = n
        kw["xyz"[a] + "Max"] = m
    ctx = G.Context(**kw)
    ctx.upload(g[f"{case}_psi0"])
    ctx.evolve(int(steps), 0)
    assert rel_l2(ctx.download(), g[f"{case}_imag"]) <= TOL64
    ctx.evolve(int(steps), 1)
    assert rel_l2(ctx.download(), g[f"{case}_real"]) <= TOL64
    ctx.close()
