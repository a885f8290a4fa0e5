"""The batched sweep's workload (64 uneven 2D FP32-bulk instances, 128^2-512^2,
g ~ U[100,2000], Omega ~ U[0,0.95] gauge, omega_y/omega_x ~ U[0.8,1.2]) against
the FP64 oracle, instance by instance, and batch == alone bit for bit.

Tolerances (FP32 bulk / FP64 reductions against the FP64 oracle): psi relative
L2 <= 1e-5; energy and chemical potential relative <= 1e-6.
"""
import numpy as np
import pytest

import bench
import gpue_amd as G
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

STEPS = 20


def psi0(ctx, seed):
    grids = [ctx.grid(0), ctx.grid(1)]
    psi = bench.seeded_gaussian(ctx.shape, grids, seed, False, 0.01)
    dV = (grids[0][1] - grids[0][0]) * (grids[1][1] - grids[1][0])
    return psi / np.sqrt(np.sum(np.abs(psi) ** 2) * dV)


def test_config4_mix_against_oracle():
    sizes, specs = bench.config4_specs()
    assert sorted(set(int(n) for n in sizes)) == [128, 192, 256, 384, 512]
    ctxs, psis = [], []
    for cfg, seed in specs:
        c = G.Context(precision=32, **cfg)
        p = psi0(c, seed)
        c.upload(p)
        ctxs.append(c)
        psis.append(p)
    b = G.Batch(ctxs)
    b.evolve(STEPS, 0)
    e_b, m_b = b.energy()
    worst = [0.0, 0.0, 0.0]
    for k, ((cfg, seed), c) in enumerate(zip(specs, ctxs)):
        orc = O.Oracle(**cfg)
        orc.set_wfc(psis[k])
        orc.evolve(STEPS, 0)
        eo, mo = orc.energy()
        e_psi = rel_l2(c.download(), orc.wfc())
        e_e, e_m = abs(e_b[k] - eo) / abs(eo), abs(m_b[k] - mo) / abs(mo)
        worst = [max(worst[0], e_psi), max(worst[1], e_e), max(worst[2], e_m)]
        assert e_psi <= 1e-5 and e_e <= 1e-6 and e_m <= 1e-6, (k, int(sizes[k]), e_psi, e_e, e_m)
    print("worst psi, E, mu:", worst)
    b.close()
    for c in ctxs:
        c.close()


def test_config4_batch_equals_alone_large_groups():
    # one instance of each large size, alone and in the full batch: bit-identical
    sizes, specs = bench.config4_specs()
    ctxs = [G.Context(precision=32, **cfg) for cfg, _ in specs]
    for c, (cfg, seed) in zip(ctxs, specs):
        c.upload(psi0(c, seed))
    b = G.Batch(ctxs)
    b.evolve(STEPS, 0)
    b.evolve(6, 1)
    for n in (256, 384, 512):
        k = int(np.nonzero(sizes == n)[0][0])
        cfg, seed = specs[k]
        a = G.Context(precision=32, **cfg)
        a.upload(psi0(a, seed))
        a.set_graph_steps(0)
        a.evolve(STEPS, 0)
        a.evolve(6, 1, first_step=STEPS)
        assert np.array_equal(a.download(), ctxs[k].download()), n
        assert a.energy() == ctxs[k].energy(), n
        a.close()
    b.close()
    for c in ctxs:
        c.close()
