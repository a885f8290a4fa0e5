"""ctypes wrapper of the CPU FP64 oracle (oracle/libgpue_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(gpue_amd) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpue_oracle.so")

V_HARMONIC, V_TABLE = 0, 1
A_ROTATION, A_ZERO, A_TABLE = 0, 1, 2
K_STANDARD, K_TABLE = 0, 1
FIELD_V, FIELD_K, FIELD_AX, FIELD_AY, FIELD_AZ = range(5)


class OrParams(ctypes.Structure):
    _fields_ = [
        ("dimnum", ctypes.c_int), ("xDim", ctypes.c_int), ("yDim", ctypes.c_int), ("zDim", ctypes.c_int),
        ("xMax", ctypes.c_double), ("yMax", ctypes.c_double), ("zMax", ctypes.c_double),
        ("hbar", ctypes.c_double), ("mass", ctypes.c_double), ("g", ctypes.c_double),
        ("omega", ctypes.c_double), ("omegaX", ctypes.c_double), ("omegaY", ctypes.c_double),
        ("omegaZ", ctypes.c_double), ("gammaY", ctypes.c_double),
        ("x0_shift", ctypes.c_double), ("y0_shift", ctypes.c_double), ("z0_shift", ctypes.c_double),
        ("gdt", ctypes.c_double), ("dt", ctypes.c_double),
        ("gpe", ctypes.c_int), ("corotating", ctypes.c_int),
        ("v_kind", ctypes.c_int), ("a_kind", ctypes.c_int), ("k_kind", ctypes.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        vp = ctypes.c_void_p
        L.or_fft_lines.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_long, ctypes.c_int]
        L.or_fft_nd.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_dft_bruteforce.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
        L.or_create.argtypes = [ctypes.POINTER(OrParams)]
        L.or_create.restype = vp
        L.or_destroy.argtypes = [vp]
        L.or_set_field.argtypes = [vp, ctypes.c_int, dp]
        L.or_rebuild.argtypes = [vp]
        L.or_get_field.argtypes = [vp, ctypes.c_int, dp]
        L.or_get_grid.argtypes = [vp, ctypes.c_int, dp]
        L.or_set_wfc.argtypes = [vp, dp]
        L.or_get_wfc.argtypes = [vp, dp]
        L.or_set_g.argtypes = [vp, ctypes.c_double]
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_max_threads.restype = ctypes.c_int
        L.or_dV.argtypes = [vp]
        L.or_dV.restype = ctypes.c_double
        L.or_step.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_set_ramp.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_ramp_factor.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.or_ramp_factor.restype = ctypes.c_double
        L.or_parsum.argtypes = [vp]
        L.or_energy.argtypes = [vp, dp]
        L.or_energy.restype = ctypes.c_double
        L.or_gauge_spectral.argtypes = [vp]
        L.or_gauge_spectral.restype = ctypes.c_double
        L.or_energy_consistent.argtypes = [vp, dp]
        L.or_energy_consistent.restype = ctypes.c_double
        L.or_evolve.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp,
                                ctypes.c_int, ip]
        L.or_evolve.restype = ctypes.c_int
        L.or_parsum_raw.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.or_find_vortex.argtypes = [ip, dp, ctypes.c_double, ctypes.c_int, dp]
        L.or_find_vortex.restype = ctypes.c_int
        L.or_vort_pos.argtypes = [ip, ctypes.c_int, ip, ctypes.c_int]
        L.or_vort_pos.restype = ctypes.c_int
        L.or_ls_fit.argtypes = [ip, ctypes.c_int, dp, ctypes.c_int, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def fft_lines(a, n, howmany, stride, dist, sign):
    """In-place unnormalised FFT of `howmany` lines of a complex128 array."""
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    lib().or_fft_lines(a.ctypes.data, n, howmany, stride, dist, sign)


def fft_nd(a, sign):
    """In-place unnormalised nD FFT of a complex128 array shaped (z, y, x) or (y, x) or (x,)."""
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    shp = a.shape[::-1] + (1,) * (3 - a.ndim)
    lib().or_fft_nd(a.ctypes.data, shp[0], shp[1], shp[2], sign)


def dft_bruteforce(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    out = np.empty_like(a)
    lib().or_dft_bruteforce(a.ctypes.data, out.ctypes.data, a.size, sign)
    return out


def parsum_raw(wfc, dimnum, dx, dy, dz):
    """GPUE parSum on a complex128 array shaped (z,y,x)/(y,x)/(x,); in place."""
    shp = wfc.shape[::-1] + (1,) * (3 - wfc.ndim)
    lib().or_parsum_raw(wfc.ctypes.data, dimnum, shp[0], shp[1], shp[2], dx, dy, dz)


def find_vortex(wfc2d, radius, x):
    """GPUE Tracker::findVortex on a square complex128 (xDim, xDim) array; returns (marker, found)."""
    wfc2d = np.ascontiguousarray(wfc2d, dtype=np.complex128)
    xDim = wfc2d.shape[1]
    marker = np.zeros(xDim * xDim, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    found = lib().or_find_vortex(_ip(marker), _dp(wfc2d.view(np.float64)), radius, xDim, _dp(x))
    return marker.reshape(xDim, xDim), found


def vort_pos(marker):
    marker = np.ascontiguousarray(marker, dtype=np.int32)
    xDim = marker.shape[1]
    n = lib().or_vort_pos(_ip(marker), xDim, _ip(np.zeros(3, np.int32)), 0)
    out = np.zeros(3 * max(n, 1), dtype=np.int32)
    lib().or_vort_pos(_ip(marker), xDim, _ip(out), n)
    return out[: 3 * n].reshape(n, 3)


def ls_fit(ijw, wfc2d):
    """GPUE Tracker::lsFit: sub-grid (xd, yd) for each (i, j, winding) row of ijw."""
    ijw = np.ascontiguousarray(ijw, dtype=np.int32)
    wfc2d = np.ascontiguousarray(wfc2d, dtype=np.complex128)
    n = ijw.shape[0]
    out = np.zeros(2 * max(n, 1))
    lib().or_ls_fit(_ip(ijw), n, _dp(wfc2d.view(np.float64)), wfc2d.shape[1], _dp(out))
    return out[: 2 * n].reshape(n, 2)


def vort_arrange(cur, prev):
    """Tracker::vortArrange (src/tracker.cu:234-256) on rows (x, xd, y, yd, winding):
    for each i, the vortex of cur[i:] nearest to prev[i] (squared distance of the
    sub-grid positions, converted to int by truncation as GPUE's int dist_t does)
    is swapped into place i; the first of equal distances wins.  GPUE reads
    prev[i] for every i < len(cur); here only i < len(prev) are arranged (the
    rest keep their order), as the drop-in does."""
    cur = [tuple(r) for r in cur]
    for i in range(min(len(cur), len(prev))):
        dist, index = 0x7FFFFFFF, i
        for j in range(i, len(cur)):
            dt = int((prev[i][1] - cur[j][1]) * (prev[i][1] - cur[j][1]) +
                     (prev[i][3] - cur[j][3]) * (prev[i][3] - cur[j][3]))
            if dist > dt:
                dist, index = dt, j
        cur[index], cur[i] = cur[i], cur[index]
    return cur


class Oracle:
    """One GPUE problem on the CPU: grids, fields, operator arrays and psi."""

    def __init__(self, **kw):
        p = OrParams()
        defaults = dict(dimnum=2, xDim=64, yDim=64, zDim=1, xMax=10.0, yMax=10.0, zMax=10.0,
                        hbar=1.0, mass=1.0, g=0.0, omega=0.0, omegaX=1.0, omegaY=1.0, omegaZ=1.0,
                        gammaY=1.0, x0_shift=0.0, y0_shift=0.0, z0_shift=0.0, gdt=1e-3, dt=1e-3,
                        gpe=0, corotating=0, v_kind=V_HARMONIC, a_kind=A_ROTATION, k_kind=K_STANDARD)
        defaults.update(kw)
        for k, v in defaults.items():
            setattr(p, k, v)
        self.params = defaults
        self.dimnum = defaults["dimnum"]
        self.shape = {1: (defaults["xDim"],), 2: (defaults["yDim"], defaults["xDim"]),
                      3: (defaults["zDim"], defaults["yDim"], defaults["xDim"])}[self.dimnum]
        self.N = int(np.prod(self.shape))
        self._h = lib().or_create(ctypes.byref(p))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_destroy(self._h)
            self._h = None

    def set_field(self, which, vals):
        vals = np.ascontiguousarray(vals, dtype=np.float64).reshape(-1)
        assert vals.size == self.N
        lib().or_set_field(self._h, which, _dp(vals))

    def rebuild(self):
        lib().or_rebuild(self._h)

    def field(self, which):
        out = np.empty(self.N)
        lib().or_get_field(self._h, which, _dp(out))
        return out.reshape(self.shape)

    def grid(self, axis):
        n = [self.params["xDim"], self.params["yDim"], self.params["zDim"]][axis % 3]
        if axis % 3 == 1 and self.dimnum < 2:
            n = 1
        if axis % 3 == 2 and self.dimnum < 3:
            n = 1
        out = np.empty(n)
        lib().or_get_grid(self._h, axis, _dp(out))
        return out

    def set_wfc(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.complex128).reshape(-1)
        assert psi.size == self.N
        lib().or_set_wfc(self._h, _dp(psi.view(np.float64)))

    def wfc(self):
        out = np.empty(self.N, dtype=np.complex128)
        lib().or_get_wfc(self._h, _dp(out.view(np.float64)))
        return out.reshape(self.shape)

    def set_g(self, g):
        lib().or_set_g(self._h, g)

    def set_ramp(self, ramp_type, num_steps):
        """-R ramp (src/evolution.cu:409-428,811-832); steps count from 0."""
        lib().or_set_ramp(self._h, int(ramp_type), int(num_steps))

    @property
    def dV(self):
        return lib().or_dV(self._h)

    def step(self, i, gstate):
        lib().or_step(self._h, i, gstate)

    def parsum(self):
        lib().or_parsum(self._h)

    def energy(self):
        mu = ctypes.c_double()
        e = lib().or_energy(self._h, ctypes.byref(mu))
        return e, mu.value

    def gauge_spectral(self):
        """hbar <A.k> of the evolved Hamiltonian (spectral, per axis)."""
        return lib().or_gauge_spectral(self._h)

    def energy_consistent(self):
        """energy_calc with the gauge term replaced by gauge_spectral()."""
        mu = ctypes.c_double()
        e = lib().or_energy_consistent(self._h, ctypes.byref(mu))
        return e, mu.value

    def evolve(self, nsteps, gstate, energy_every=0, thr=-1.0, max_energies=100000):
        en = np.zeros(max(1, max_energies))
        ne = ctypes.c_int()
        done = lib().or_evolve(self._h, nsteps, gstate, energy_every, thr, _dp(en), max_energies,
                               ctypes.byref(ne))
        return done, en[: min(ne.value, max_energies)]


def set_threads(n):
    """OpenMP threads of the oracle's following calls."""
    lib().or_set_threads(int(n))


def max_threads():
    return int(lib().or_max_threads())


def ramp_factor(ramp_type, omega, i, num_steps):
    """omega_0 of step i (src/evolution.cu:412-427)."""
    return lib().or_ramp_factor(int(ramp_type), float(omega), int(i), int(num_steps))
