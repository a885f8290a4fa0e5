"""The oracle's step, parSum and energy pinned by the reference's KATs and by
closed forms (src/unit_test.cu:841-947, 1162-1278)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2, seeded_psi


def test_parsum_kat_2d_3d():
    # src/unit_test.cu:841-947: 2+2i everywhere, dx = dy = dz = 0.1
    w = np.full((64, 64), 2 + 2j)
    O.parsum_raw(w, 2, 0.1, 0.1, 0.1)
    assert np.all(w.real == 2 / np.sqrt(32768.0 * 0.1 * 0.1))
    assert np.all(w.imag == 2 / np.sqrt(32768.0 * 0.1 * 0.1))
    w = np.full((16, 16, 16), 2 + 2j)
    O.parsum_raw(w, 3, 0.1, 0.1, 0.1)
    assert np.all(w.real == 2 / np.sqrt(32768.0 * 0.1 * 0.1 * 0.1))


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_ho_ground_state_energy(dim):
    # evolve_test (src/unit_test.cu:1162-1278): E = d * hbar omega / 2 within 1 %;
    # here in oscillator units, and tighter: the linear ground state is a Gaussian.
    n = 32 if dim == 3 else 64
    o = O.Oracle(dimnum=dim, xDim=n, yDim=n, zDim=n, xMax=7, yMax=7, zMax=7, gdt=1e-3, dt=1e-3, omegaZ=1.0)
    grids = [o.grid(a) for a in range(dim)]
    o.set_wfc(seeded_psi(o.shape, grids, seed=dim))
    o.parsum()
    o.evolve(4000 if dim < 3 else 2500, 0)
    e, mu = o.energy()
    assert abs(e - dim / 2) < 1e-5 * dim
    assert abs(mu - dim / 2) < 1e-5 * dim
    # real time: the stationary state keeps its energy and norm
    o.evolve(2000, 1)
    e2, _ = o.energy()
    psi = o.wfc()
    assert abs(e2 - e) < 1e-9
    assert abs(np.sum(np.abs(psi) ** 2) * o.dV - 1) < 1e-12


def test_plane_wave_kinetic_energy():
    # K = hbar^2 k^2 / 2m for a grid plane wave (energy_calc kinetic term, src/split_op.cu:256-273)
    n, L = 64, 8.0
    o = O.Oracle(dimnum=2, xDim=n, yDim=n, xMax=L, yMax=L, hbar=1.3, mass=0.7)
    x = o.grid(0)
    kx = 5 * np.pi / L
    ky = -3 * np.pi / L
    psi = np.exp(1j * (kx * x[None, :] + ky * x[:, None]))
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * o.dV)
    o.set_wfc(psi)
    V = o.field(O.FIELD_V)
    e, _ = o.energy()
    expect = 1.3 ** 2 * (kx ** 2 + ky ** 2) / (2 * 0.7) + np.sum(V) * o.dV / (4 * L * L) * 1.0
    assert abs(e - expect) < 1e-10 * abs(expect)


def test_gauge_energy_forward_difference():
    # gauge term of energy_calc: Re conj(psi) i hbar (Ax Dx psi + Ay Dy psi), D = forward
    # difference on the flat array (src/split_op.cu:312-360, derive src/kernels.cu:58-71).
    # For psi = exp(i k x) the interior value is -hbar Ax sin(k dx)/dx |psi|^2.
    n, L, om = 32, 6.0, 0.4
    o = O.Oracle(dimnum=2, xDim=n, yDim=n, xMax=L, yMax=L, corotating=1, omega=om)
    o2 = O.Oracle(dimnum=2, xDim=n, yDim=n, xMax=L, yMax=L, corotating=0, omega=om)
    x = o.grid(0)
    dx = x[1] - x[0]
    k = 4 * np.pi / (2 * L)
    psi = np.exp(1j * k * x)[None, :] * np.ones((n, 1))
    o.set_wfc(psi)
    o2.set_wfc(psi)
    Ax = o.field(O.FIELD_AX)
    Ay = o.field(O.FIELD_AY)
    # closed form per point: hbar * Re[conj(psi) * i * (Ax Dx psi + Ay Dy psi)]
    dxpsi = (np.exp(1j * k * dx) - 1) / dx * psi
    dxpsi[:, -1] = (np.roll(psi.reshape(-1), -1).reshape(n, n)[:, -1] - psi[:, -1]) / dx
    dxpsi[-1, -1] = psi[-1, -1] / dx
    dypsi = np.zeros_like(psi)
    dypsi[-1, :] = psi[-1, :] / dx
    term = np.real(np.conj(psi) * 1j * (Ax * dxpsi + Ay * dypsi))
    e_g, _ = o.energy()
    e_0, _ = o2.energy()
    assert abs((e_g - e_0) - np.sum(term) * o.dV) < 1e-10 * max(1.0, abs(e_g))


def test_thomas_fermi_bound():
    # 2D TF limit for V = r^2/2: mu_TF = sqrt(g/pi), E_TF = 2/3 mu_TF (sanity bound)
    g = 2000.0
    o = O.Oracle(dimnum=2, xDim=128, yDim=128, xMax=12, yMax=12, g=g, gpe=1, gdt=2e-3)
    grids = [o.grid(0), o.grid(1)]
    o.set_wfc(seeded_psi(o.shape, grids, seed=2))
    o.parsum()
    o.evolve(3000, 0)
    e, mu = o.energy()
    mu_tf = np.sqrt(g / np.pi)
    assert 0.97 * mu_tf < mu < 1.06 * mu_tf
    assert 0.97 * (2 / 3) * mu_tf < e < 1.06 * (2 / 3) * mu_tf


def test_matlab_gpe2d_crosscheck(golden):
    # matlab/GPE_2d.m imaginary-time loop (lines 87-110) transcribed independently in numpy
    g = golden("gpe2d_matlab.npz")
    n = int(g["n"])
    o = O.Oracle(dimnum=2, xDim=n, yDim=n, xMax=float(g["xmax"]), yMax=float(g["xmax"]), g=float(g["g"]), gpe=1,
                 gdt=float(g["Dt"]))
    o.set_wfc(g["psi0"])
    o.step(0, 0)
    assert rel_l2(o.wfc(), g["psi_1"]) < 1e-13
    for i in range(1, int(g["steps"])):
        o.step(i, 0)
    assert rel_l2(o.wfc(), g["psi_%d" % int(g["steps"])]) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_gauge_order_crosscheck(golden, dim):
    # GPUE's step order with alternating per-axis gauge operators, transcribed in numpy
    g = golden(f"gpue_gauge_{dim}d.npz")
    n = int(g["n"])
    L = float(g["xmax"])
    o = O.Oracle(dimnum=dim, xDim=n, yDim=n, zDim=n, xMax=L, yMax=L, zMax=L, g=float(g["g"]), gpe=1,
                 corotating=1, omega=float(g["omega"]), omegaZ=2.0, gdt=float(g["gdt"]), dt=float(g["dt"]))
    o.set_wfc(g["psi0"])
    steps = int(g["steps"])
    for i in range(steps):
        o.step(i, 0)
    assert rel_l2(o.wfc(), g["psi_imag"]) < 1e-12
    for i in range(steps):
        o.step(i, 1)
    assert rel_l2(o.wfc(), g["psi_real"]) < 1e-12


def test_real_time_unitarity_and_energy():
    o = O.Oracle(dimnum=2, xDim=64, yDim=64, xMax=8, yMax=8, g=80, gpe=1, corotating=1, omega=0.5, dt=5e-4)
    grids = [o.grid(0), o.grid(1)]
    o.set_wfc(seeded_psi(o.shape, grids, seed=5, phase=False))
    o.parsum()
    e0, _ = o.energy()
    o.evolve(400, 1)
    psi = o.wfc()
    assert abs(np.sum(np.abs(psi) ** 2) * o.dV - 1) < 1e-12
    e1, _ = o.energy()
    assert abs(e1 - e0) < 2e-2 * abs(e0)


def test_vortex_fixtures(golden):
    d = golden("vortex_cases.npz")
    for c in ["v1", "v2", "v3"]:
        m, found = O.find_vortex(d[c + "_psi"], float(d[c + "_radius"]), d["x"])
        assert np.array_equal(m, d[c + "_marker"]), c
        assert found == int(np.count_nonzero(d[c + "_marker"])), c
    lst = O.vort_pos(d["v1_marker"])
    assert [tuple(r) for r in lst] == [(20, 20, 1), (20, 40, 1), (40, 30, 1), (45, 45, 1)]


def test_early_stop_rule():
    # src/evolution.cu:899-902: stop when |E_old - E| < thr * E_old in imaginary time
    o = O.Oracle(dimnum=2, xDim=32, yDim=32, xMax=6, yMax=6, gdt=1e-2)
    grids = [o.grid(0), o.grid(1)]
    o.set_wfc(seeded_psi(o.shape, grids, seed=9))
    o.parsum()
    done, en = o.evolve(5000, 0, energy_every=10, thr=1e-6)
    assert done < 5000 and done % 10 == 1
    assert abs(en[-1] - en[-2]) < 1e-6 * en[-2]
    assert abs(en[-2] - en[-3]) >= 1e-6 * en[-3]


def test_ls_fit_linear_vortex_closed_form():
    # a field that is exactly linear in the grid indices, psi = (i - i0) + 1j (j - j0),
    # is fitted exactly by the least-squares plane: the core is (i0, j0)
    n = 16
    I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    for (i0, j0) in [(5.3, 7.8), (10.01, 2.99), (8.5, 8.5)]:
        psi = (I - i0) + 1j * (J - j0)
        ijw = np.array([[int(i0), int(j0), 1]], dtype=np.int32)
        xy = O.ls_fit(ijw, psi)
        assert abs(xy[0, 0] - i0) < 1e-12 and abs(xy[0, 1] - j0) < 1e-12


def test_ramp_factors_and_cumulative_exponent():
    # src/evolution.cu:409-428: omega_0 of step i; the operators are raised to
    # it every step, so after step i the exponent carries the running product,
    # which telescopes to omega/(i+1) (type 1) or omega*(i+1)/n (type 2)
    om, n = 0.7, 9
    for t in (1, 2):
        prod = 1.0
        for i in range(n):
            f = O.ramp_factor(t, om, i, n)
            expect = (om if i == 0 else i / (i + 1)) if t == 1 else (om / n if i == 0 else (i + 1) / i)
            assert f == expect
            prod *= f
            closed = om / (i + 1) if t == 1 else om * (i + 1) / n
            assert abs(prod - closed) <= 4e-15 * closed


def test_ramp_imaginary_time_is_exponent_scaling():
    # imaginary time: raising the real positive operator exp(-pA gdt) to the
    # step factors one after another depends only on their product, i.e. the
    # ramp scales the gauge exponent.  Two ramp types whose running products
    # agree at step 1 (type 1: omega * 1/2; type 2, n = 4: omega/4 * 2) through
    # different pow sequences must give the same step.
    kw = dict(dimnum=1, xDim=64, xMax=8.0, g=20.0, gpe=1, corotating=1, omega=0.5, gdt=1e-3)
    a = O.Oracle(**kw)
    b = O.Oracle(**kw)
    rng = np.random.default_rng(2)
    psi = np.exp(-a.grid(0) ** 2 / 2) * np.exp(1j * rng.uniform(0, 2 * np.pi, 64))
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * a.dV)
    a.set_ramp(1, 4)
    b.set_ramp(2, 4)
    # step 0 (factors omega and omega/4) advances the operators; step 1 is then
    # taken from the same state
    a.set_wfc(psi)
    b.set_wfc(psi)
    a.step(0, 0)
    b.step(0, 0)
    a.set_wfc(psi)
    b.set_wfc(psi)
    a.step(1, 0)
    b.step(1, 0)
    assert np.linalg.norm(a.wfc() - b.wfc()) <= 1e-13 * np.linalg.norm(b.wfc())


def _vortex_gaussian(o, kappa=0.3):
    X, Y = np.meshgrid(o.grid(0), o.grid(1))
    return np.exp(-(X ** 2 + Y ** 2) / 2) * (X + 1j * Y) * np.exp(1j * kappa * X)


@pytest.mark.parametrize("hbar,om,ox,oy", [(1.0, 0.7, 1.0, 1.0), (2.5, 0.9, 1.3, 0.6)])
def test_gauge_spectral_angular_momentum_closed_form(hbar, om, ox, oy):
    # Rotation gauge Ax = -y om ox, Ay = x om oy (src/operators.cu:630-647); the evolved
    # term hbar <A.k> = hbar om (ox <-y kx> + oy <x ky>).  For psi = (x + iy) exp(-r^2/2)
    # exp(i kappa x) (unnormalised), <x ky> = <-y kx> = pi/2 (the boost adds <-y kappa> = 0),
    # so the term is hbar om (ox + oy) pi / 2.  The spectral derivative and the Riemann sum of
    # a Gaussian are exact to rounding on this box.
    o = O.Oracle(dimnum=2, xDim=128, yDim=128, xMax=8.0, yMax=8.0, corotating=1, omega=om,
                 omegaX=ox, omegaY=oy, hbar=hbar)
    o.set_wfc(_vortex_gaussian(o))
    want = hbar * om * (ox + oy) * np.pi / 2
    assert abs(o.gauge_spectral() - want) <= 1e-12 * want


def test_gauge_spectral_against_forward_difference():
    # GPUE's energy_calc gauge term (+i hbar A.grad, forward differences) is minus the
    # evolved term up to the difference error: the sum of the two falls as dx^2 (the
    # first-order error cancels on this field), and the consistent energy is GPUE's with
    # the sign of its gauge term flipped, to that error.
    gaps = []
    for n in (64, 128):
        o = O.Oracle(dimnum=2, xDim=n, yDim=n, xMax=8.0, yMax=8.0, corotating=1, omega=0.7)
        o0 = O.Oracle(dimnum=2, xDim=n, yDim=n, xMax=8.0, yMax=8.0, corotating=0, omega=0.7)
        psi = _vortex_gaussian(o)
        o.set_wfc(psi)
        o0.set_wfc(psi)
        e_gpue, mu_gpue = o.energy()
        e_none, mu_none = o0.energy()
        e_cons, mu_cons = o.energy_consistent()
        spec = o.gauge_spectral()
        assert abs(e_cons - (e_none + spec)) <= 1e-12 * abs(e_cons)
        assert abs(mu_cons - (mu_none + spec)) <= 1e-12 * abs(mu_cons)
        gaps.append(abs((e_gpue - e_none) + spec) / spec)
    assert gaps[0] < 0.03 and 3.5 < gaps[0] / gaps[1] < 4.5, gaps


def test_energy_consistent_is_conserved_in_real_time():
    # In an anisotropic rotating trap L_z is not conserved, so GPUE's energy (gauge term of
    # the wrong sign) drifts while the energy of the evolved Hamiltonian stays put.
    o = O.Oracle(dimnum=2, xDim=64, yDim=64, xMax=8.0, yMax=8.0, corotating=1, omega=0.5,
                 omegaX=1.0, omegaY=1.6, dt=1e-3, gpe=0)
    X, Y = np.meshgrid(o.grid(0), o.grid(1))
    o.set_wfc(np.exp(-((X - 0.5) ** 2 + Y ** 2) / 2) * np.exp(0.4j * Y))
    o.parsum()
    e0, _ = o.energy_consistent()
    g0, _ = o.energy()
    o.evolve(400, 1)
    e1, _ = o.energy_consistent()
    g1, _ = o.energy()
    assert abs(e1 - e0) < 1e-4 * abs(e0)
    assert abs(g1 - g0) > 100 * abs(e1 - e0)


@pytest.mark.parametrize("case", ["line", "rect", "box"])
def test_gauge_box_crosscheck(golden, case):
    # GPUE's step with per-axis gauge operators on a 1D line and on boxes whose axes differ,
    # transcribed in numpy (tests/golden/make_golden.py, gpue_gauge_box)
    g = golden("gpue_gauge_box.npz")
    dims = [int(v) for v in g[f"{case}_dims"]]
    maxes = [float(v) for v in g[f"{case}_maxes"]]
    gg, om, gam, oz, steps, gdt, dt = [float(v) for v in g[f"{case}_par"]]
    kw = dict(dimnum=len(dims), g=gg, gpe=1, corotating=1, omega=om, gammaY=gam, omegaZ=oz, gdt=gdt, dt=dt)
    for a, (n, m) in enumerate(zip(dims, maxes)):
        kw["xyz"[a] + "Dim"] = n
        kw["xyz"[a] + "Max"] = m
    o = O.Oracle(**kw)
    o.set_wfc(g[f"{case}_psi0"])
    for i in range(int(steps)):
        o.step(i, 0)
    assert rel_l2(o.wfc(), g[f"{case}_imag"]) < 1e-12
    for i in range(int(steps)):
        o.step(i, 1)
    assert rel_l2(o.wfc(), g[f"{case}_real"]) < 1e-12


def test_energy_and_mu_weights_by_parseval():
    # E = <K> + <V> + g/2 int|psi|^4 (energy_calc, src/split_op.cu:208-375) and the chemical
    # potential mu = <psi|H_GP|psi> = <K> + <V> + g int|psi|^4, both computed here with
    # numpy.fft on a random state: pins the interaction weights g/2 and g and the kinetic
    # normalisation independently of the oracle's own sums
    n, L, g, hbar, mass = 64, 8.0, 137.0, 1.1, 0.9
    o = O.Oracle(dimnum=2, xDim=n, yDim=n, xMax=L, yMax=L, g=g, gpe=1, hbar=hbar, mass=mass)
    x = o.grid(0)
    dx = x[1] - x[0]
    rng = np.random.default_rng(5)
    psi = (np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / 3) * (1 + 0.3 * rng.uniform(-1, 1, (n, n)))
           * np.exp(1j * rng.uniform(0, 2 * np.pi, (n, n))))
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * dx * dx)
    o.set_wfc(psi)
    k = 2 * np.pi * np.fft.fftfreq(n, d=dx)
    K = hbar ** 2 * (k[None, :] ** 2 + k[:, None] ** 2) / (2 * mass)
    ekin = np.sum(K * np.abs(np.fft.fft2(psi)) ** 2) / (n * n) * dx * dx
    V = 0.5 * mass * (x[None, :] ** 2 + x[:, None] ** 2)
    epot = np.sum(V * np.abs(psi) ** 2) * dx * dx
    e4 = np.sum(np.abs(psi) ** 4) * dx * dx
    e, mu = o.energy()
    assert abs(e - (ekin + epot + 0.5 * g * e4)) < 1e-11 * abs(e)
    assert abs(mu - (ekin + epot + g * e4)) < 1e-11 * abs(mu)
