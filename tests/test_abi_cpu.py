"""The C ABI library loads and exports every entry point include/gpue_amd.h
declares; host-side validation and defaults work without a GPU."""
import ctypes
import os
import re

import pytest

import gpue_amd as G

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "gpue_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpue_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = G.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(G.ABI_SYMBOLS) == syms


def test_params_default_matches_gpue_parser():
    # src/parser.cu:14-67 and src/init.cu:98-137 (atoms = 1, box_size < 0)
    p = G.params_default()
    assert (p.dimnum, p.xDim, p.yDim, p.zDim) == (2, 256, 256, 256)
    assert (p.gdt, p.dt, p.omegaX, p.omegaY, p.omegaZ, p.gammaY) == (1e-4, 1e-4, 6.283, 6.283, 6.283, 1.0)
    assert p.hbar == 1.05457148e-34 and p.mass == 1.4431607e-25
    HBAR, PI, m, a_s = 1.05457148e-34, 3.141592653589793, 1.4431607e-25, 4.76e-9
    g = 4 * HBAR * HBAR * PI * (a_s / m) * (m * 6.283 / (2 * PI * HBAR)) ** 0.5
    assert abs(p.g - g) <= 1e-15 * g
    Rxy = 15 ** 0.2 * (a_s * (m * 6.283 / HBAR) ** 0.5) ** 0.2
    assert abs(p.xMax - 6 * Rxy * (HBAR / (2 * m * 6.283)) ** 0.5) <= 1e-15 * p.xMax


@pytest.mark.parametrize("kw,status", [
    (dict(dimnum=4), G.GPUE_EINVAL),
    (dict(precision=16), G.GPUE_EINVAL),
    (dict(xDim=100, yDim=100), G.GPUE_EUNSUPPORTED),
    (dict(xDim=8192, yDim=8192), G.GPUE_EUNSUPPORTED),
    (dict(xMax=-1.0), G.GPUE_EINVAL),
    (dict(hbar=0.0), G.GPUE_EINVAL),
])
def test_create_rejects_before_touching_the_device(kw, status):
    p = G.make_params(**kw)
    h = ctypes.c_void_p()
    assert G.lib().gpue_create(ctypes.byref(p), ctypes.byref(h)) == status
    assert not h.value


def test_null_arguments():
    lib = G.lib()
    assert lib.gpue_create(None, None) == G.GPUE_EINVAL
    assert lib.gpue_energy(None, None, None) == G.GPUE_EINVAL
    assert lib.gpue_evolve(None, 0, 1, 0, 0, 0.0, None, 0, None) == G.GPUE_EINVAL


def test_eqn_parse_rejects_bad_buffers():
    # host-only entry point: checks its buffer arguments before writing any
    lib = G.lib()
    nodes = (G.EqnNode * 4)()
    nn, td = ctypes.c_int(), ctypes.c_int()
    assert lib.gpue_eqn_parse(b"x*x", None, None, 0, nodes, -1, ctypes.byref(nn), ctypes.byref(td)) == G.GPUE_EINVAL
    names = (ctypes.c_char_p * 1)(b"a")
    assert lib.gpue_eqn_parse(b"a*x", names, None, 1, None, 0, ctypes.byref(nn), ctypes.byref(td)) == G.GPUE_EINVAL
    assert lib.gpue_eqn_parse(b"x*x", None, None, -1, None, 0, ctypes.byref(nn), ctypes.byref(td)) == G.GPUE_EINVAL
    # a buffer too small gets what fits and GPUE_EINVAL; a large enough one the whole program
    assert lib.gpue_eqn_parse(b"x*x+y", None, None, 0, None, 0, ctypes.byref(nn), ctypes.byref(td)) == G.GPUE_OK
    need = nn.value
    assert need > 1
    small = (G.EqnNode * 1)()
    assert lib.gpue_eqn_parse(b"x*x+y", None, None, 0, small, 1, ctypes.byref(nn), ctypes.byref(td)) == G.GPUE_EINVAL
    big = (G.EqnNode * need)()
    assert lib.gpue_eqn_parse(b"x*x+y", None, None, 0, big, need, ctypes.byref(nn), ctypes.byref(td)) == G.GPUE_OK
    assert lib.gpue_eqn_eval(big, need, 2.0, 3.0, 0.0, 0.0) == 7.0
    assert lib.gpue_step_bytes(None, 0, None, None) == G.GPUE_EINVAL


GPUE_BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpue_amd", "bin", "gpue")


@pytest.mark.parametrize("args,message", [
    (["-P", "5"], "unsupported option -P"),            # GPUE's laser power (src/parser.cu:295-299)
    (["-q", "1"], "unsupported option -q"),
    (["--precision", "16"], "--precision must be 64 or 32"),
    (["--precison", "32"], "unsupported option --precison"),
])
def test_cli_rejects_flags_outside_the_step(tmp_path, args, message):
    # parse_args fails before any device call, so this runs without a GPU
    import subprocess

    r = subprocess.run([GPUE_BIN, "-d", str(tmp_path)] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert message in r.stderr


def test_sweep_units_match_their_generator():
    # one translation unit per (precision, axis length, plan), as gen_units.py writes them
    import importlib.util
    import os

    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpue_amd", "csrc")
    spec = importlib.util.spec_from_file_location("gen_units", os.path.join(here, "gen_units.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    units = sorted(f for f in os.listdir(here) if f.startswith("sweep_f") and f.endswith(".hip"))
    want = sorted(f"sweep_{p}_{n}_v{v}.hip" for n in gen.SIZES for p in ("f64", "f32") for v in (0, 1))
    assert units == want
    for f in units:
        prec, n, v = f[:-4].split("_")[1:]
        T = "double" if prec == "f64" else "float"
        text = open(os.path.join(here, f)).read()
        assert f"GPUE_SWEEP_TU({T}, {n}, {v[1:]}, {f[:-4]})" in text, f
    # every generated length is a supported grid size of the library (the dispatch lists the same)
    src = open(os.path.join(here, "sweep_dispatch.hip")).read()
    for n in gen.SIZES:
        assert f"GPUE_CASE({n})" in src and f"case {n}:" in src, n
