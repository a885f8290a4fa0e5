"""gpue_amd -- MI355X (gfx950) split-operator step for GPUE, host-side mirror.

The compute path is libgpue_amd.so (hand-written HIP kernels behind the C ABI
declared in include/gpue_amd.h).  This module binds that ABI with ctypes and
mirrors the pieces of GPUE's host interface that drive the step:

* ``Params`` / ``params_default``  -- GPUE's Grid keys for the step
  (include/ds.h:86-197, defaults src/parser.cu:14-67),
* ``Context``                        -- one problem: fields, psi, ``evolve``
  (src/evolution.cu:260-908), ``parsum`` (src/split_op.cu:83-129),
  ``energy`` (src/split_op.cu:208-375), ``find_vortex``
  (src/tracker.cu:102-161),
* ``eqn_parse`` / ``eqn_eval``        -- GPUE equation strings
  (src/dynamic.cu:187-603).

There is no fallback: if the shared library is missing or cannot be loaded,
importing the binding raises, so nothing silently runs elsewhere.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpue_amd.so")

GPUE_OK, GPUE_EINVAL, GPUE_ENOMEM, GPUE_EUNSUPPORTED, GPUE_EHIP, GPUE_ENONFINITE, GPUE_EBUDGET, GPUE_EPARSE, GPUE_ETIMEOUT = range(9)
STATUS_NAMES = ["OK", "EINVAL", "ENOMEM", "EUNSUPPORTED", "EHIP", "ENONFINITE", "EBUDGET", "EPARSE", "ETIMEOUT"]
FIELD_V, FIELD_K, FIELD_AX, FIELD_AY, FIELD_AZ = range(5)
A_ROTATION, A_CONSTANT = 0, 1

# symbols declared in include/gpue_amd.h
ABI_SYMBOLS = [
    "gpue_params_default", "gpue_create", "gpue_destroy", "gpue_last_error", "gpue_set_field",
    "gpue_set_field_eqn", "gpue_upload_wfc", "gpue_download_wfc", "gpue_evolve", "gpue_parsum",
    "gpue_energy", "gpue_energy_consistent", "gpue_vortex_marker_2d", "gpue_vortex_marker_2d_noalloc", "gpue_vortex_list_2d", "gpue_get_grid", "gpue_set_g",
    "gpue_sync", "gpue_set_ramp", "gpue_autotune", "gpue_set_variants", "gpue_get_variants", "gpue_set_graph_steps", "gpue_prepare_evolve", "gpue_set_timing", "gpue_timing_get", "gpue_step_bytes", "gpue_sweep_profile", "gpue_math_eval", "gpue_eqn_parse", "gpue_eqn_eval",
    "gpue_batch_create", "gpue_batch_set_active", "gpue_batch_set_graph_mode", "gpue_batch_autotune", "gpue_batch_evolve", "gpue_batch_energy", "gpue_batch_last_error",
    "gpue_batch_destroy",
    "gpue_server_create", "gpue_server_submit", "gpue_server_poll", "gpue_server_drain", "gpue_server_clock",
    "gpue_server_destroy",
]


class GpueError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"GPUE_{name}: {msg}")


class Params(ctypes.Structure):
    _fields_ = [
        ("dimnum", ctypes.c_int), ("xDim", ctypes.c_int), ("yDim", ctypes.c_int), ("zDim", ctypes.c_int),
        ("xMax", ctypes.c_double), ("yMax", ctypes.c_double), ("zMax", ctypes.c_double),
        ("hbar", ctypes.c_double), ("mass", ctypes.c_double), ("g", ctypes.c_double),
        ("omega", ctypes.c_double), ("omegaX", ctypes.c_double), ("omegaY", ctypes.c_double),
        ("omegaZ", ctypes.c_double), ("gammaY", ctypes.c_double),
        ("x0_shift", ctypes.c_double), ("y0_shift", ctypes.c_double), ("z0_shift", ctypes.c_double),
        ("gdt", ctypes.c_double), ("dt", ctypes.c_double),
        ("gpe", ctypes.c_int), ("corotating", ctypes.c_int), ("a_kind", ctypes.c_int),
        ("precision", ctypes.c_int), ("device", ctypes.c_int),
    ]


class EvolveResult(ctypes.Structure):
    _fields_ = [("steps_done", ctypes.c_int), ("early_stop", ctypes.c_int), ("n_energies", ctypes.c_int),
                ("energy", ctypes.c_double), ("mu", ctypes.c_double)]


class EqnNode(ctypes.Structure):
    _fields_ = [("val", ctypes.c_double), ("is_dynamic", ctypes.c_int), ("var", ctypes.c_char),
                ("left", ctypes.c_int), ("right", ctypes.c_int), ("op_num", ctypes.c_int)]


_lib = None


def lib():
    """Load libgpue_amd.so (raises if it is missing: there is no other backend)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip, cp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p
        PP = ctypes.POINTER(Params)
        L.gpue_params_default.argtypes = [PP]
        L.gpue_create.argtypes = [PP, ctypes.POINTER(vp)]
        L.gpue_destroy.argtypes = [vp]
        L.gpue_destroy.restype = None
        L.gpue_last_error.argtypes = [vp]
        L.gpue_last_error.restype = cp
        L.gpue_set_field.argtypes = [vp, ctypes.c_int, dp]
        L.gpue_set_field_eqn.argtypes = [vp, ctypes.c_int, cp, ctypes.POINTER(cp), dp, ctypes.c_int]
        L.gpue_upload_wfc.argtypes = [vp, dp]
        L.gpue_download_wfc.argtypes = [vp, dp]
        L.gpue_evolve.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp,
                                  ctypes.c_int, ctypes.POINTER(EvolveResult)]
        L.gpue_parsum.argtypes = [vp]
        L.gpue_energy.argtypes = [vp, dp, dp]
        L.gpue_energy_consistent.argtypes = [vp, dp, dp]
        L.gpue_vortex_marker_2d.argtypes = [vp, ctypes.c_double, ip, ip]
        L.gpue_vortex_marker_2d_noalloc.argtypes = [vp, ctypes.c_double, ip, ip]
        L.gpue_vortex_list_2d.argtypes = [vp, ctypes.c_double, ip, ctypes.c_int, ip]
        L.gpue_get_grid.argtypes = [vp, ctypes.c_int, dp]
        L.gpue_set_g.argtypes = [vp, ctypes.c_double]
        L.gpue_sync.argtypes = [vp]
        L.gpue_set_ramp.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.gpue_autotune.argtypes = [vp, ctypes.c_int, ip]
        L.gpue_set_variants.argtypes = [vp, ip]
        L.gpue_set_graph_steps.argtypes = [vp, ctypes.c_int]
        L.gpue_prepare_evolve.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.gpue_set_timing.argtypes = [vp, ctypes.c_int]
        L.gpue_timing_get.argtypes = [vp, ctypes.c_int, ctypes.POINTER(cp), ctypes.POINTER(ctypes.c_long), dp]
        L.gpue_step_bytes.argtypes = [vp, ctypes.c_int, dp, dp]
        L.gpue_sweep_profile.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ctypes.POINTER(cp), dp, dp]
        L.gpue_math_eval.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_long]
        L.gpue_eqn_parse.argtypes = [cp, ctypes.POINTER(cp), dp, ctypes.c_int, ctypes.POINTER(EqnNode),
                                     ctypes.c_int, ip, ip]
        L.gpue_eqn_eval.argtypes = [ctypes.POINTER(EqnNode), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double]
        L.gpue_eqn_eval.restype = ctypes.c_double
        L.gpue_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
        L.gpue_batch_set_active.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.gpue_batch_set_graph_mode.argtypes = [vp, ctypes.c_int]
        L.gpue_get_variants.argtypes = [vp, ip]
        L.gpue_batch_autotune.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, ip]
        L.gpue_batch_evolve.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.gpue_batch_energy.argtypes = [vp, dp, dp]
        L.gpue_batch_last_error.argtypes = [vp]
        L.gpue_batch_last_error.restype = cp
        L.gpue_batch_destroy.argtypes = [vp]
        L.gpue_batch_destroy.restype = None
        L.gpue_server_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(vp)]
        L.gpue_server_submit.argtypes = [vp, ctypes.c_void_p, ctypes.c_ulonglong]
        L.gpue_server_poll.argtypes = [vp, ctypes.c_void_p, ctypes.c_int]
        L.gpue_server_drain.argtypes = [vp, ctypes.c_double]
        L.gpue_server_clock.argtypes = [vp]
        L.gpue_server_clock.restype = ctypes.c_double
        L.gpue_server_destroy.argtypes = [vp]
        L.gpue_server_destroy.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def params_default():
    """GPUE's parser.cu defaults (SI units, 2D 256^2)."""
    p = Params()
    lib().gpue_params_default(ctypes.byref(p))
    return p


def make_params(**kw):
    """Params in harmonic-oscillator units (hbar = m = 1) unless overridden."""
    p = params_default()
    base = dict(dimnum=2, xDim=256, yDim=256, zDim=1, xMax=10.0, yMax=10.0, zMax=10.0, hbar=1.0, mass=1.0,
                g=0.0, omega=0.0, omegaX=1.0, omegaY=1.0, omegaZ=1.0, gammaY=1.0, x0_shift=0.0, y0_shift=0.0,
                z0_shift=0.0, gdt=1e-3, dt=1e-3, gpe=0, corotating=0, a_kind=A_ROTATION, precision=64, device=0)
    base.update(kw)
    for k, v in base.items():
        setattr(p, k, v)
    return p


def math_eval(fn, x, device=0):
    """Device exp / sin / cos of the fused operators: "exp", "sin", "cos" (FP64),
    "sin32", "cos32" (FP32, input rounded to float) and the short forms for
    |x| <= 1/32, "exp_small", "sin_small", "cos_small"."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    st = lib().gpue_math_eval(device, {"exp": 0, "sin": 1, "cos": 2, "sin32": 3, "cos32": 4, "exp_small": 5, "sin_small": 6,
                                          "cos_small": 7}[fn], _dp(x), _dp(y),
                              x.size)
    if st != GPUE_OK:
        raise GpueError(st, "gpue_math_eval failed")
    return y


def _names(mapping):
    mapping = mapping or {}
    keys = list(mapping.keys())
    arr_t = ctypes.c_char_p * max(1, len(keys))
    names = arr_t(*[k.encode() for k in keys])
    vals = np.array([float(mapping[k]) for k in keys] or [0.0])
    return names, vals, len(keys)


def eqn_parse(expr, doubles=None):
    """GPUE parse_eqn + tree_to_array: returns (list of node dicts, time_dependent)."""
    names, vals, n = _names(doubles)
    nn = ctypes.c_int()
    td = ctypes.c_int()
    st = lib().gpue_eqn_parse(expr.encode(), names, _dp(vals), n, None, 0, ctypes.byref(nn), ctypes.byref(td))
    if st != GPUE_OK:
        raise GpueError(st, f"cannot parse {expr!r}")
    nodes = (EqnNode * max(1, nn.value))()
    st = lib().gpue_eqn_parse(expr.encode(), names, _dp(vals), n, nodes, nn.value, ctypes.byref(nn), ctypes.byref(td))
    if st != GPUE_OK:
        raise GpueError(st, f"cannot parse {expr!r}")
    return nodes, bool(td.value)


def eqn_eval(nodes, x=0.0, y=0.0, z=0.0, t=0.0):
    return lib().gpue_eqn_eval(nodes, len(nodes), x, y, z, t)


class Context:
    """One GPUE problem on the GPU (wraps gpue_ctx)."""

    def __init__(self, params=None, **kw):
        self._h = None
        self.params = params if params is not None else make_params(**kw)
        p = self.params
        self.dimnum = p.dimnum
        dims = (p.xDim, p.yDim if p.dimnum > 1 else 1, p.zDim if p.dimnum > 2 else 1)
        self.shape = {1: (dims[0],), 2: (dims[1], dims[0]), 3: (dims[2], dims[1], dims[0])}[p.dimnum]
        self.N = int(np.prod(self.shape))
        h = ctypes.c_void_p()
        st = lib().gpue_create(ctypes.byref(p), ctypes.byref(h))
        if st != GPUE_OK:
            raise GpueError(st, "gpue_create failed")
        self._h = h

    def close(self):
        if self._h:
            lib().gpue_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st, what):
        if st != GPUE_OK:
            msg = lib().gpue_last_error(self._h)
            raise GpueError(st, f"{what}: {msg.decode() if msg else ''}")

    # fields -----------------------------------------------------------------
    def set_field(self, which, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        assert v.size == self.N
        self._check(lib().gpue_set_field(self._h, which, _dp(v)), "set_field")

    def set_field_eqn(self, which, expr, doubles=None):
        names, vals, n = _names(doubles)
        self._check(lib().gpue_set_field_eqn(self._h, which, expr.encode(), names, _dp(vals), n), "set_field_eqn")

    def grid(self, axis):
        n = [self.params.xDim, self.params.yDim, self.params.zDim][axis % 3]
        if (axis % 3) >= self.dimnum:
            n = 1
        out = np.empty(n)
        self._check(lib().gpue_get_grid(self._h, axis, _dp(out)), "get_grid")
        return out

    def set_g(self, g):
        self._check(lib().gpue_set_g(self._h, g), "set_g")

    # psi --------------------------------------------------------------------
    def upload(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.complex128).reshape(-1)
        assert psi.size == self.N
        self._check(lib().gpue_upload_wfc(self._h, _dp(psi.view(np.float64))), "upload_wfc")

    def download(self):
        out = np.empty(self.N, dtype=np.complex128)
        self._check(lib().gpue_download_wfc(self._h, _dp(out.view(np.float64))), "download_wfc")
        return out.reshape(self.shape)

    # step ---------------------------------------------------------------------
    def evolve(self, nsteps, gstate, first_step=0, energy_every=0, thr=-1.0, max_energies=4096):
        en = np.zeros(max(1, max_energies))
        r = EvolveResult()
        self._check(lib().gpue_evolve(self._h, first_step, nsteps, gstate, energy_every, thr, _dp(en), max_energies,
                                      ctypes.byref(r)), "evolve")
        return r, en[: min(r.n_energies, max_energies)]

    def parsum(self):
        self._check(lib().gpue_parsum(self._h), "parsum")

    def energy(self):
        e = ctypes.c_double()
        m = ctypes.c_double()
        self._check(lib().gpue_energy(self._h, ctypes.byref(e), ctypes.byref(m)), "energy")
        return e.value, m.value

    def energy_consistent(self):
        """(energy, mu) with the gauge term of the evolved Hamiltonian, hbar <A.k>
        (spectral), in place of GPUE's forward-difference +i hbar A.grad."""
        e = ctypes.c_double()
        m = ctypes.c_double()
        self._check(lib().gpue_energy_consistent(self._h, ctypes.byref(e), ctypes.byref(m)), "energy_consistent")
        return e.value, m.value

    def vortex_marker(self, radius):
        n = self.params.xDim
        marker = np.zeros(n * n, dtype=np.int32)
        found = ctypes.c_int()
        self._check(lib().gpue_vortex_marker_2d(self._h, radius, _ip(marker), ctypes.byref(found)), "vortex")
        return marker.reshape(n, n), found.value

    def vortex_list(self, radius, max_vortices=100000):
        out = np.zeros(3 * max_vortices, dtype=np.int32)
        n = ctypes.c_int()
        self._check(lib().gpue_vortex_list_2d(self._h, radius, _ip(out), max_vortices, ctypes.byref(n)), "vortex")
        return out[: 3 * min(n.value, max_vortices)].reshape(-1, 3)

    def sync(self):
        self._check(lib().gpue_sync(self._h), "sync")

    def autotune(self, reps=5):
        v = np.zeros(3, dtype=np.int32)
        self._check(lib().gpue_autotune(self._h, reps, _ip(v)), "autotune")
        return tuple(int(a) for a in v)

    def set_graph_steps(self, steps):
        self._check(lib().gpue_set_graph_steps(self._h, steps), "set_graph_steps")

    def prepare_evolve(self, nsteps, gstate, first_step=0, energy_every=0):
        """Capture the step blocks evolve(nsteps, gstate, first_step, energy_every) replays."""
        self._check(lib().gpue_prepare_evolve(self._h, first_step, nsteps, gstate, energy_every), "prepare_evolve")

    def sweep_profile(self, gstate, reps=50):
        """Per-kernel cost of the step: [{name, us, per_step}] from back-to-back
        graph replays of each distinct sweep (in place on fresh copies of psi, copies subtracted), HIP events on the
        context stream."""
        m = 32
        names = (ctypes.c_char_p * m)()
        us = np.zeros(m)
        per = np.zeros(m)
        n = ctypes.c_int()
        self._check(lib().gpue_sweep_profile(self._h, gstate, reps, m, ctypes.byref(n), names, _dp(us), _dp(per)),
                    "sweep_profile")
        return [dict(name=names[k].decode(), us=float(us[k]), per_step=float(per[k])) for k in range(min(n.value, m))]

    def set_ramp(self, ramp_type, num_steps):
        """GPUE's -R ramp of the gauge exponent (0 off, 1 or 2) over an evolve
        call of num_steps steps."""
        self._check(lib().gpue_set_ramp(self._h, int(ramp_type), int(num_steps)), "set_ramp")

    def variants(self):
        v = np.zeros(3, dtype=np.int32)
        self._check(lib().gpue_get_variants(self._h, _ip(v)), "get_variants")
        return tuple(int(a) for a in v)

    def set_variants(self, variants):
        v = np.ascontiguousarray(variants, dtype=np.int32)
        self._check(lib().gpue_set_variants(self._h, _ip(v)), "set_variants")

    # measurement --------------------------------------------------------------
    def set_timing(self, on):
        self._check(lib().gpue_set_timing(self._h, 1 if on else 0), "set_timing")

    def timing(self):
        out = {}
        k = 0
        while True:
            name = ctypes.c_char_p()
            n = ctypes.c_long()
            ms = ctypes.c_double()
            if lib().gpue_timing_get(self._h, k, ctypes.byref(name), ctypes.byref(n), ctypes.byref(ms)) != GPUE_OK:
                break
            out[name.value.decode()] = (n.value, ms.value)
            k += 1
        return out

    def step_bytes(self, gstate):
        b = ctypes.c_double()
        s = ctypes.c_double()
        self._check(lib().gpue_step_bytes(self._h, gstate, ctypes.byref(b), ctypes.byref(s)), "step_bytes")
        return b.value, s.value


class Batch:
    """Many independent contexts advanced together (gpue_batch): one launch per
    sweep covers every instance of a shape; each shape group replays two steps
    from a graph on its own stream."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c._h.value for c in self.contexts])
        h = ctypes.c_void_p()
        st = lib().gpue_batch_create(arr, len(self.contexts), ctypes.byref(h))
        if st != GPUE_OK:
            raise GpueError(st, "gpue_batch_create failed")
        self._h = h

    def _check(self, st, what):
        if st != GPUE_OK:
            msg = lib().gpue_batch_last_error(self._h)
            raise GpueError(st, f"{what}: {msg.decode() if msg else ''}")

    def autotune(self, gstate=0, steps=20):
        """Variants per shape group and the graph mode, timed on the whole batch;
        states are restored.  Returns (best microseconds per batch step, graph mode)."""
        us = ctypes.c_double()
        mode = ctypes.c_int()
        self._check(lib().gpue_batch_autotune(self._h, gstate, steps, ctypes.byref(us), ctypes.byref(mode)),
                    "batch_autotune")
        return us.value, mode.value

    def set_graph_mode(self, mode):
        """0: one graph per shape group (default); 1: one graph for the whole batch; 2: no graphs."""
        self._check(lib().gpue_batch_set_graph_mode(self._h, int(mode)), "set_graph_mode")

    def set_active(self, index, active):
        self._check(lib().gpue_batch_set_active(self._h, index, 1 if active else 0), "batch_set_active")

    def evolve(self, nsteps, gstate):
        self._check(lib().gpue_batch_evolve(self._h, nsteps, gstate), "batch_evolve")

    def energy(self):
        """(energy, mu) arrays of every instance, evaluated concurrently."""
        e = np.zeros(len(self.contexts))
        m = np.zeros(len(self.contexts))
        self._check(lib().gpue_batch_energy(self._h, _dp(e), _dp(m)), "batch_energy")
        return e, m

    def close(self):
        if getattr(self, "_h", None):
            lib().gpue_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Request(ctypes.Structure):
    _fields_ = [
        ("params", Params), ("seed", ctypes.c_ulonglong), ("random_phase", ctypes.c_int),
        ("check_every", ctypes.c_int), ("max_imag_steps", ctypes.c_int), ("energy_tol", ctypes.c_double),
        ("real_steps", ctypes.c_int), ("quench_factor", ctypes.c_double),
        ("psi_out", ctypes.POINTER(ctypes.c_double)),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_ulonglong), ("status", ctypes.c_int), ("imag_steps", ctypes.c_int),
        ("real_steps", ctypes.c_int), ("energy", ctypes.c_double), ("mu", ctypes.c_double),
        ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
        ("t_dev_start", ctypes.c_double), ("t_dev_done", ctypes.c_double),
    ]


class Server:
    """Continuous-batching solve service (gpue_server): requests are admitted in
    arrival order while a slot and the HBM budget allow; each admitted request
    runs on its own stream, overlapping with the others."""

    def __init__(self, budget_bytes, max_slots, device=0):
        h = ctypes.c_void_p()
        st = lib().gpue_server_create(device, float(budget_bytes), int(max_slots), ctypes.byref(h))
        if st != GPUE_OK:
            raise GpueError(st, "gpue_server_create failed")
        self._h = h
        self._keep = {}   # id -> psi_out array, alive until its result is polled

    def submit(self, rid, params, seed=1, random_phase=False, check_every=100, max_imag_steps=1000,
               energy_tol=1e-8, real_steps=0, quench_factor=1.0, psi_out=None):
        r = Request()
        r.params = params
        r.seed, r.random_phase = int(seed), 1 if random_phase else 0
        r.check_every, r.max_imag_steps, r.energy_tol = int(check_every), int(max_imag_steps), float(energy_tol)
        r.real_steps, r.quench_factor = int(real_steps), float(quench_factor)
        if psi_out is not None:
            assert psi_out.dtype == np.complex128 and psi_out.flags.c_contiguous
            r.psi_out = psi_out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            self._keep[int(rid)] = psi_out
        st = lib().gpue_server_submit(self._h, ctypes.byref(r), int(rid))
        if st != GPUE_OK:
            self._keep.pop(int(rid), None)
            raise GpueError(st, "gpue_server_submit failed")

    def poll(self, max_results=1024):
        buf = (Result * max_results)()
        n = lib().gpue_server_poll(self._h, buf, max_results)
        out = []
        for k in range(n):
            r = buf[k]
            self._keep.pop(int(r.id), None)
            out.append(dict(id=int(r.id), status=int(r.status), imag_steps=int(r.imag_steps),
                            real_steps=int(r.real_steps), energy=float(r.energy), mu=float(r.mu),
                            t_submit=float(r.t_submit), t_start=float(r.t_start), t_done=float(r.t_done),
                            t_dev_start=float(r.t_dev_start), t_dev_done=float(r.t_dev_done)))
        return out

    def drain(self, timeout_s=600.0):
        st = lib().gpue_server_drain(self._h, float(timeout_s))
        if st != GPUE_OK:
            raise GpueError(st, "gpue_server_drain timed out" if st == GPUE_ETIMEOUT else "gpue_server_drain")

    def clock(self):
        return float(lib().gpue_server_clock(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().gpue_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
