"""The `gpue` executable's input side against the oracle: a wavefunction read
with -r (src/fileIO.cu:13-33, src/init.cu:525-539), equation fields from an -I
parameter file (src/dynamic.cu:187-222: continuation lines, and V gets
+0.5*mass*(Ax*Ax+Ay*Ay+Az*Az) when it is not the file's last entry), gauge
fields from -A file (src/operators.cu:261-270, 505-520: A = omega * value), and
the vortex output over two print points (vortArrange, src/tracker.cu:234-256,
and vLoc_, src/evolution.cu:564-579)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2, seeded_psi

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPUE = os.path.join(ROOT, "gpue_amd", "bin", "gpue")
N, L, G_INT = 64, 8.0, 60.0


def read_wfc(d, name, step, n):
    re = np.loadtxt(os.path.join(d, f"{name}_{step}"))
    im = np.loadtxt(os.path.join(d, f"{name}i_{step}"))
    return (re + 1j * im).reshape(n, n)


def write_wfc(d, psi):
    flat = psi.reshape(-1)
    np.savetxt(os.path.join(d, "wfc_load"), flat.real, fmt="%.16e")
    np.savetxt(os.path.join(d, "wfci_load"), flat.imag, fmt="%.16e")


def grid_xy():
    x = np.array([-L + i * (2 * L / N) for i in range(N)])
    return x, np.meshgrid(x, x)   # X[y, x], Y[y, x]


def base_cmd(d, extra):
    return [GPUE, "-d", d, "-x", str(N), "-y", str(N), "-b", str(L), "-B", "1", "-M", "1", "-X", "1", "-Y", "1",
            "-N", str(G_INT), "-s", "-T", "1e-3", "-t", "5e-4", "-W", "-r"] + extra


def psi0(seed, phase):
    x, _ = grid_xy()
    p = seeded_psi((N, N), [x, x], seed=seed, phase=phase)
    return p / np.sqrt(np.sum(np.abs(p) ** 2) * (2 * L / N) ** 2)


def test_read_wfc_and_param_file_fields(tmp_path):
    d = str(tmp_path)
    p0 = psi0(11, True)
    write_wfc(d, p0)
    with open(os.path.join(d, "param.cfg"), "w") as f:
        f.write("Ax = 0.5*y\n")
        f.write("V = 0.5*(x*x +\n")          # continued on the next line (no '=')
        f.write("   y*y)\n")
        f.write("Ay = 0.5*x\n")              # V is not last: GPUE appends 1/2 m |A|^2 to it
    out = subprocess.run(base_cmd(d, ["-l", "-w", "0.5", "-g", "40", "-e", "0", "-p", "20", "-I", "param.cfg"]),
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr + out.stdout
    # the read-in state is used as it is (GPUE normalises only its own initial state)
    assert rel_l2(read_wfc(d, "wfc_0_const", 0, N), p0) < 1e-15
    _, (X, Y) = grid_xy()
    Ax, Ay = 0.5 * Y, 0.5 * X
    V = 0.5 * (X * X + Y * Y) + 0.5 * 1.0 * (Ax * Ax + Ay * Ay)
    orc = O.Oracle(dimnum=2, xDim=N, yDim=N, xMax=L, yMax=L, g=G_INT, gpe=1, corotating=1, omega=0.5, gdt=1e-3)
    for which, f in ((O.FIELD_AX, Ax), (O.FIELD_AY, Ay), (O.FIELD_V, V)):
        orc.set_field(which, f)
    orc.rebuild()
    orc.set_wfc(p0)
    for i in range(20):
        orc.step(i, 0)
    assert rel_l2(read_wfc(d, "wfc_0_const", 20, N), orc.wfc()) < 1e-10
    # without the appended term the state differs (the quirk is exercised)
    orc2 = O.Oracle(dimnum=2, xDim=N, yDim=N, xMax=L, yMax=L, g=G_INT, gpe=1, corotating=1, omega=0.5, gdt=1e-3)
    for which, f in ((O.FIELD_AX, Ax), (O.FIELD_AY, Ay), (O.FIELD_V, 0.5 * (X * X + Y * Y))):
        orc2.set_field(which, f)
    orc2.rebuild()
    orc2.set_wfc(p0)
    for i in range(20):
        orc2.step(i, 0)
    assert rel_l2(orc2.wfc(), orc.wfc()) > 1e-6


def test_gauge_fields_from_files(tmp_path):
    d = str(tmp_path)
    p0 = psi0(12, True)
    write_wfc(d, p0)
    _, (X, Y) = grid_xy()
    fx = -Y + 0.05 * X * X
    fy = X + 0.02 * Y
    np.savetxt(os.path.join(d, "Axgauge"), fx.reshape(-1), fmt="%.16e")
    np.savetxt(os.path.join(d, "Aygauge"), fy.reshape(-1), fmt="%.16e")
    om = 0.7
    out = subprocess.run(base_cmd(d, ["-l", "-w", str(om), "-A", "file", "-g", "40", "-e", "0", "-p", "20"]),
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr + out.stdout
    orc = O.Oracle(dimnum=2, xDim=N, yDim=N, xMax=L, yMax=L, g=G_INT, gpe=1, corotating=1, omega=om, gdt=1e-3)
    orc.set_field(O.FIELD_AX, om * fx)
    orc.set_field(O.FIELD_AY, om * fy)
    orc.rebuild()
    orc.set_wfc(p0)
    for i in range(20):
        orc.step(i, 0)
    assert rel_l2(read_wfc(d, "wfc_0_const", 20, N), orc.wfc()) < 1e-10


def test_vortex_arrangement_over_print_points(tmp_path):
    d = str(tmp_path)
    p0 = psi0(13, True)   # a random phase: many phase singularities
    write_wfc(d, p0)
    ps = 10
    out = subprocess.run(base_cmd(d, ["-l", "-w", "0.6", "-g", "2", "-e", str(2 * ps), "-p", str(ps), "-m", "6",
                                      "-o"]), capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr + out.stdout
    x, _ = grid_xy()
    lists = {}
    for i in (0, ps):
        w = read_wfc(d, "wfc_ev", i, N)
        m, found = O.find_vortex(w, 6.0, x)
        pos = O.vort_pos(m)
        fit = O.ls_fit(pos, w)
        lists[i] = [(int(a), float(xd), int(b), float(yd), int(wd)) for (a, b, wd), (xd, yd) in zip(pos, fit)]
        if i == ps:
            vloc = np.loadtxt(os.path.join(d, f"vLoc__{i}"), dtype=np.int64)
            assert np.array_equal(vloc, m.reshape(-1))
    assert len(lists[0]) > 3 and len(lists[ps]) > 3
    expect = O.vort_arrange(lists[ps], lists[0])
    rows = [r for r in open(os.path.join(d, f"vort_arr_{ps}")).read().splitlines()[1:] if r]
    assert len(rows) == len(expect)
    for r, (a, xd, b, yd, wd) in zip(rows, expect):
        f = r.split(",")
        assert (int(f[0]), int(f[2]), int(f[4])) == (a, b, wd)
        assert abs(float(f[1]) - xd) <= 1e-5 * max(1.0, abs(xd)) and abs(float(f[3]) - yd) <= 1e-5 * max(1.0, abs(yd))
    # the arrangement is not the raster order here (the test is not vacuous)
    assert [(a, b) for a, _, b, _, _ in expect] != [(a, b) for a, _, b, _, _ in lists[ps]]
