"""The oracle's own FFT (replacing cuFFT Z2Z, src/init.cu:309-319) pinned by a
brute-force long-double DFT, numpy.fft golden vectors and Parseval."""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 12, 16, 30, 31, 64])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_vs_bruteforce(n, sign):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = a.copy()
    O.fft_lines(b, n, 1, 1, n, sign)
    ref = O.dft_bruteforce(a, sign)
    assert np.max(np.abs(b - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_fft_golden_vectors(golden):
    g = golden("fft_vectors.npz")
    for key in g.files:
        if not key.startswith("in_") or key == "in_nd":
            continue
        n = int(key[3:])
        a = g[key]
        f = a.copy()
        O.fft_lines(f, n, 1, 1, n, -1)
        i = a.copy()
        O.fft_lines(i, n, 1, 1, n, +1)
        scale = np.max(np.abs(g[f"fwd_{n}"]))
        assert np.max(np.abs(f - g[f"fwd_{n}"])) <= 1e-13 * scale, n
        assert np.max(np.abs(i - g[f"inv_{n}"])) <= 1e-13 * scale, n
    a = g["in_nd"].copy()
    O.fft_nd(a, -1)
    assert np.max(np.abs(a - g["fwd_nd"])) <= 1e-12 * np.max(np.abs(g["fwd_nd"]))


@pytest.mark.parametrize("n", [2048, 4096])
def test_fft_long_lines_vs_numpy(n):
    # the longest lines the HIP path takes: pinned to numpy.fft (pocketfft)
    rng = np.random.default_rng(n)
    a = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    for sign, ref in ((-1, np.fft.fft(a, axis=1)), (1, np.fft.ifft(a, axis=1) * n)):
        b = a.copy()
        O.fft_lines(b, n, 3, 1, n, sign)
        assert np.max(np.abs(b - ref)) <= 1e-12 * np.max(np.abs(ref)), (n, sign)


def test_parseval_and_roundtrip():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((16, 24, 32)) + 1j * rng.standard_normal((16, 24, 32))
    f = a.copy()
    O.fft_nd(f, -1)
    assert abs(np.sum(np.abs(f) ** 2) / a.size - np.sum(np.abs(a) ** 2)) <= 1e-12 * np.sum(np.abs(a) ** 2)
    O.fft_nd(f, +1)
    assert np.max(np.abs(f / a.size - a)) <= 1e-14


def test_strided_lines_match_axis_transform():
    # y-axis lines of an x-fastest array (GPUE plan_other2d, src/ds.cu:11-29)
    rng = np.random.default_rng(4)
    a = rng.standard_normal((32, 16)) + 1j * rng.standard_normal((32, 16))
    b = a.copy()
    O.fft_lines(b, 32, 16, 16, 1, -1)
    assert np.max(np.abs(b - np.fft.fft(a, axis=0))) <= 1e-12
