"""Device exp / sin / cos of the fused operators (Cody-Waite reduction +
Taylor polynomials) against the host libm through numpy.

Tolerances: FP64 exp relative error <= 4e-16 on [-700, 700]; FP64 sin and cos
absolute error <= 4e-16 for |x| <= 100 and <= 2e-15 up to 1e6 (beyond 2^20 the
device uses the library); FP32 sin and cos absolute error <= 3e-7 against the
FP64 values at the float-rounded input; special values propagate (NaN stays
NaN, exp(-inf) = 0).
"""
import numpy as np
import pytest

import gpue_amd as G

pytestmark = pytest.mark.gpu


def test_exp_matches_libm():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-700, 700, 200000), rng.uniform(-1, 1, 100000),
                        np.array([0.0, -0.0, 1e-300, -1e-300, 0.34657359, -0.34657359, 709.5, -744.0])])
    y = G.math_eval("exp", x)
    ref = np.exp(x)
    ok = ref > 1e-300
    assert np.max(np.abs(y[ok] - ref[ok]) / ref[ok]) <= 4e-16
    s = G.math_eval("exp", np.array([np.nan, -np.inf, np.inf, -800.0, 800.0]))
    assert np.isnan(s[0]) and s[1] == 0.0 and np.isinf(s[2]) and s[3] == 0.0 and np.isinf(s[4])


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_sincos_matches_libm(fn):
    rng = np.random.default_rng(4)
    small = np.concatenate([rng.uniform(-100, 100, 200000), np.arange(-64, 65) * (np.pi / 4)])
    big = rng.uniform(-1e6, 1e6, 100000)
    ref_f = getattr(np, fn)
    assert np.max(np.abs(G.math_eval(fn, small) - ref_f(small))) <= 4e-16
    assert np.max(np.abs(G.math_eval(fn, big) - ref_f(big))) <= 2e-15
    huge = np.array([3e6, -7.5e9, 1e300])
    assert np.max(np.abs(G.math_eval(fn, huge) - ref_f(huge))) <= 1e-15
    assert np.isnan(G.math_eval(fn, np.array([np.nan, np.inf]))).all()


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_sincos_fp32_matches_libm(fn):
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-50, 50, 200000), rng.uniform(-7000, 7000, 50000),
                        np.arange(-64, 65) * (np.pi / 4)]).astype(np.float32).astype(np.float64)
    y = G.math_eval(fn + "32", x)
    assert np.max(np.abs(y - getattr(np, fn)(x))) <= 3e-7
    assert np.isnan(G.math_eval(fn + "32", np.array([np.nan, np.inf]))).all()


@pytest.mark.parametrize("fn", ["exp", "sin", "cos"])
def test_small_argument_forms_match_libm(fn):
    # the short polynomials the FP64 phase factors take when every argument of a lane
    # is within 1/32 (no reduction; truncation below 1e-18 relative)
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.uniform(-1 / 32, 1 / 32, 200000), np.array([0.0, -0.0, 1 / 32, -1 / 32, 1e-300])])
    y = G.math_eval(fn + "_small", x)
    ref = getattr(np, fn)(x)
    if fn == "exp":
        assert np.max(np.abs(y - ref) / ref) <= 4e-16
    else:
        assert np.max(np.abs(y - ref)) <= 2e-16
