"""Error paths of the ABI forced from the host (no kernel is made to fail)."""
import numpy as np
import pytest

import gpue_amd as G

pytestmark = pytest.mark.gpu


def test_nonfinite_norm_is_reported():
    with G.Context(dimnum=2, xDim=32, yDim=32, xMax=4, yMax=4) as c:
        c.upload(np.zeros((32, 32), complex))
        with pytest.raises(G.GpueError) as ei:
            c.evolve(2, 0)
        assert ei.value.status == G.GPUE_ENONFINITE


def test_bad_arguments():
    with G.Context(dimnum=2, xDim=32, yDim=32, xMax=4, yMax=4) as c:
        with pytest.raises(G.GpueError):
            c.evolve(1, 7)
        with pytest.raises(G.GpueError):
            c.set_field(9, np.zeros(32 * 32))
