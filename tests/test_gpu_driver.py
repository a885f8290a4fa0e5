"""The `gpue` executable (GPUE's CLI, files and evolve() control flow over the
C ABI) against the oracle: dumps, energy files, the one-step-stale imag->real
hand-off (src/evolution.cu:430,906; src/init.cu:470) and the vortex output."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPUE = os.path.join(ROOT, "gpue_amd", "bin", "gpue")


def read_wfc(d, name, step, n):
    re = np.loadtxt(os.path.join(d, f"{name}_{step}"))
    im = np.loadtxt(os.path.join(d, f"{name}i_{step}"))
    return (re + 1j * im).reshape(n, n)


def test_driver_matches_oracle(tmp_path):
    n, L, g, om = 64, 8.0, 60.0, 0.6
    gsteps, esteps, ps = 200, 100, 100
    d = str(tmp_path)
    cmd = [GPUE, "-d", d, "-x", str(n), "-y", str(n), "-b", str(L), "-B", "1", "-M", "1", "-X", "1", "-Y", "1",
           "-N", str(g), "-s", "-l", "-w", str(om), "-g", str(gsteps), "-e", str(esteps), "-p", str(ps),
           "-T", "1e-3", "-t", "5e-4", "-W", "-m", "5.0", "-E", "1e-15", "50"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr + out.stdout
    # initial state: kstd_wfc with GPUE's sqrt(sum |psi| dV) normalisation (src/init.cu:231-307)
    psi0 = read_wfc(d, "WFC", 0, n)
    orc = O.Oracle(dimnum=2, xDim=n, yDim=n, xMax=L, yMax=L, g=g, gpe=1, corotating=1, omega=om, gdt=1e-3, dt=5e-4)
    x = orc.grid(0)
    Rxy = 15 ** 0.2 * (1 * 4.76e-9 * np.sqrt(1 * 1.0 / 1)) ** 0.2
    a0 = np.sqrt(1 / (2 * 1 * 1.0))
    amp = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / (Rxy * Rxy * a0 * a0))
    dx = x[1] - x[0]
    expect0 = amp / np.sqrt(np.sum(np.abs(amp)) * dx * dx)
    assert rel_l2(psi0, expect0) < 1e-14
    # imaginary time dumps and energies
    orc.set_wfc(psi0)
    energies = []
    for i in range(gsteps):
        if i % ps == 0:
            assert rel_l2(read_wfc(d, "wfc_0_const", i, n), orc.wfc()) < 1e-10
        if i == gsteps - 1:
            stale = orc.wfc().copy()        # host copy at the start of the last step
        orc.step(i, 0)
        if i % 50 == 0:
            energies.append(orc.energy()[0])
    ei = np.loadtxt(os.path.join(d, "energyi.dat"))
    assert np.allclose(ei, energies, rtol=1e-5)       # written with ostream's 6 digits
    # real time starts from the stale copy (GPUE's hand-off)
    orc.set_wfc(stale)
    for i in range(esteps):
        if i % ps == 0:
            got = read_wfc(d, "wfc_ev", i, n)
            assert rel_l2(got, orc.wfc()) < 1e-10
            # vortex output: findVortex + vortPos + lsFit of the same state
            m, found = O.find_vortex(got, 5.0, x)
            lst = O.vort_pos(m)
            fit = O.ls_fit(lst, got)
            rows = [l for l in open(os.path.join(d, f"vort_arr_{i}")).read().splitlines()[1:] if l]
            assert len(rows) == len(lst)
            if i == 0:
                for r, (a, b, w), (xd, yd) in zip(rows, lst, fit):
                    f = r.split(",")
                    assert (int(f[0]), int(f[2]), int(f[4])) == (a, b, w)
                    assert abs(float(f[1]) - xd) < 1e-5 * max(1, abs(xd)) and abs(float(f[3]) - yd) < 1e-5 * max(1, abs(yd))
        orc.step(i, 1)
    params = open(os.path.join(d, "Params.dat")).read()
    assert params.startswith("[Params]\n") and "xMax=8" in params


def test_driver_ramp_matches_oracle(tmp_path):
    # -R 2: ramped imaginary time, GPUE's wfc_0_ramp file names (src/evolution.cu:451-453)
    n, L, g, om, gsteps, ps = 64, 8.0, 60.0, 0.6, 40, 20
    d = str(tmp_path)
    cmd = [GPUE, "-d", d, "-x", str(n), "-y", str(n), "-b", str(L), "-B", "1", "-M", "1", "-X", "1", "-Y", "1",
           "-N", str(g), "-s", "-l", "-w", str(om), "-g", str(gsteps), "-e", "0", "-p", str(ps), "-T", "1e-3",
           "-W", "-R", "2", "-o"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr + out.stdout
    orc = O.Oracle(dimnum=2, xDim=n, yDim=n, xMax=L, yMax=L, g=g, gpe=1, corotating=1, omega=om, gdt=1e-3, dt=1e-4)
    orc.set_ramp(2, gsteps)
    orc.set_wfc(read_wfc(d, "WFC", 0, n))
    for i in range(gsteps):
        if i % ps == 0:
            assert rel_l2(read_wfc(d, "wfc_0_ramp", i, n), orc.wfc()) < 1e-10
        orc.step(i, 0)


def test_driver_rejects_unsupported_flags(tmp_path):
    out = subprocess.run([GPUE, "-d", str(tmp_path), "-K", "1"], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0 and "unsupported option" in out.stderr
    out = subprocess.run([GPUE, "-d", str(tmp_path), "-x", "100", "-y", "100"], capture_output=True, text=True,
                         timeout=60)
    assert out.returncode != 0 and "gpue_create" in out.stderr
