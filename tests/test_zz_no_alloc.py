"""gpue_vortex_marker_2d_noalloc (the allocation-free entry point: its scratch
is a fixed device array, no hipMalloc) against gpue_vortex_marker_2d on the same
states: identical markers and counts, on two calls in a row, the second on a
smaller grid, and a rejected grid above the scratch's maximum."""
import ctypes

import numpy as np
import pytest

import gpue_amd as G
from tests.conftest import seeded_psi

pytestmark = pytest.mark.gpu


def _noalloc(ctx, radius):
    n = ctx.params.xDim
    m = np.zeros(n * n, dtype=np.int32)
    found = ctypes.c_int()
    st = G.lib().gpue_vortex_marker_2d_noalloc(ctx._h, radius, m.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                ctypes.byref(found))
    return st, m.reshape(n, n), found.value


def _state(n, seed):
    kw = dict(dimnum=2, xDim=n, yDim=n, xMax=10, yMax=10, g=300, gpe=1, corotating=1, omega=0.8, gdt=1e-3, dt=1e-4)
    c = G.Context(**kw)
    c.upload(seeded_psi(c.shape, [c.grid(0), c.grid(1)], seed=seed, phase=True, noise=0.0))
    c.parsum()
    c.evolve(40, 0)
    return c


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_vortex_marker_matches_allocating_entry_point(prec):
    for n, seed, radius in ((256, 3, 6.0), (128, 4, 5.0)):   # the second call on a smaller grid
        c = _state(n, seed)
        if prec == 32:   # the same state in an FP32 context
            w = c.download()
            c.close()
            c = G.Context(precision=32, dimnum=2, xDim=n, yDim=n, xMax=10, yMax=10)
            c.upload(w)
        st, m, found = _noalloc(c, radius)
        assert st == G.GPUE_OK
        ref, ref_found = c.vortex_marker(radius)
        assert found == ref_found and found > 0
        assert np.array_equal(m, ref)
        c.close()


def test_noalloc_rejects_grids_above_its_scratch():
    with G.Context(dimnum=2, xDim=1024, yDim=1024, xMax=10, yMax=10) as ok:
        assert _noalloc(ok, 3.0)[0] == G.GPUE_OK
    with G.Context(dimnum=2, xDim=64, yDim=32, xMax=10, yMax=10) as rect:
        assert _noalloc(rect, 3.0)[0] == G.GPUE_EUNSUPPORTED
