"""Batched instances: every instance of a batch must evolve exactly as it does
alone (same kernels per instance; bit-identical), including uneven sizes,
mixed step parities, inactive instances and both precisions."""
import numpy as np
import pytest

import gpue_amd as G
from tests.conftest import seeded_psi

pytestmark = pytest.mark.gpu


def make(n, g, om, prec=64, seed=1, gamma=1.0):
    kw = dict(dimnum=2, xDim=n, yDim=n, xMax=10, yMax=10, g=g, gpe=1, corotating=1, omega=om, gammaY=gamma,
              gdt=1e-3, dt=5e-4, precision=prec)
    c = G.Context(**kw)
    psi = seeded_psi(c.shape, [c.grid(0), c.grid(1)], seed=seed, phase=True)
    c.upload(psi)
    c.parsum()
    c.set_graph_steps(0)
    return c, kw, psi


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("prec", [64, 32])
def test_batch_equals_alone(prec, mode):
    # five shape groups: more than the batch's four streams, so the two with the
    # least work (64^2 and 48^2) share one
    specs = [(64, 100.0, 0.3), (128, 300.0, 0.7), (64, 150.0, 0.5), (192, 500.0, 0.9), (128, 50.0, 0.0),
             (96, 200.0, 0.4), (48, 80.0, 0.2)]
    batch_ctx, alone = [], []
    for k, (n, g, om) in enumerate(specs):
        c, kw, psi = make(n, g, om, prec, seed=k + 1, gamma=1.0 + 0.05 * k)
        batch_ctx.append(c)
        a = G.Context(**kw)
        a.upload(psi)
        a.parsum()
        a.set_graph_steps(0)
        alone.append(a)
    # give instance 2 an odd step count so the batch mixes parities
    batch_ctx[2].evolve(1, 0)
    alone[2].evolve(1, 0)
    b = G.Batch(batch_ctx)
    b.set_graph_mode(mode)   # 1: one graph for the whole batch (a branch per shape group); 2: no graphs
    b.evolve(10, 0)
    b.evolve(7, 0)
    b.set_active(4, False)
    b.evolve(6, 0)
    for k, (c, a) in enumerate(zip(batch_ctx, alone)):
        first = 1 if k == 2 else 0
        steps = 23 if k != 4 else 17
        a.evolve(steps, 0, first_step=first)
        assert np.array_equal(c.download(), a.download()), k
    b.evolve(9, 1)
    for k, (c, a) in enumerate(zip(batch_ctx, alone)):
        if k == 4:
            continue
        first = (1 if k == 2 else 0) + 23
        a.evolve(9, 1, first_step=first)
        assert np.array_equal(c.download(), a.download()), k
        assert c.energy() == a.energy()
    b.close()


def test_batch_energy_equals_per_context():
    ctxs = []
    for k, (n, g, om) in enumerate([(64, 100.0, 0.3), (128, 300.0, 0.7), (64, 150.0, 0.0)]):
        c, _, _ = make(n, g, om, 64 if k != 1 else 32, seed=k + 3)
        ctxs.append(c)
    b = G.Batch(ctxs)
    b.evolve(12, 0)
    # the first evaluation captures each instance's energy launches as a graph,
    # the later ones replay it; a parameter change must drop the captured graph
    for rnd in range(4):
        e, m = b.energy()
        for k, c in enumerate(ctxs):
            assert (e[k], m[k]) == c.energy(), (rnd, k)
        if rnd == 1:
            ctxs[0].set_g(120.0)
        b.evolve(6, 0)
    b.close()


def test_batch_autotune_restores_state_and_keeps_parity():
    # the batch's own tuner times variants per shape group and the graph mode on
    # the instances themselves, then restores psi, step counters and pending
    # norms; the tuned batch then advances within the FP64 bar of lone contexts
    specs = [(64, 100.0, 0.3), (128, 300.0, 0.7), (64, 150.0, 0.5), (128, 50.0, 0.0)]
    ctxs, alone = [], []
    for k, (n, g, om) in enumerate(specs):
        c, kw, psi = make(n, g, om, 64, seed=k + 7)
        ctxs.append(c)
        a = G.Context(**kw)
        a.upload(psi)
        a.parsum()
        a.set_graph_steps(0)
        alone.append(a)
    b = G.Batch(ctxs)
    b.evolve(3, 0)   # a pending norm and an odd step count before tuning
    for a in alone:
        a.evolve(3, 0)
    before = [c.download() for c in ctxs]
    us, mode = b.autotune(0, 4)
    assert us > 0 and mode in (0, 1)
    for c, p in zip(ctxs, before):
        assert np.array_equal(c.download(), p)
    b.evolve(8, 0)
    b.evolve(4, 1)
    for k, (c, a) in enumerate(zip(ctxs, alone)):
        a.evolve(8, 0, first_step=3)
        a.evolve(4, 1, first_step=11)
        x, y = c.download(), a.download()
        assert np.linalg.norm(x - y) <= 1e-10 * np.linalg.norm(y), k
    b.close()
