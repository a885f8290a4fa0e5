"""Solve service (gpue_server): continuous batching with HBM-budget admission.

Every request runs on its own context and stream with deterministic kernels and
reductions, so its result must be bit-identical alone and in any mix (the
tolerance stated here is exact equality of psi, energy and step counts).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = pytest.importorskip("gpue_amd")


def _p(n, dimnum=2, prec=64, **kw):
    base = dict(dimnum=dimnum, xDim=n, yDim=n if dimnum > 1 else 1, zDim=n if dimnum > 2 else 1,
                xMax=10.0, yMax=10.0, zMax=10.0, g=200.0, gpe=1, gdt=1e-3, dt=1e-4, precision=prec)
    base.update(kw)
    return G.make_params(**base)


def _shape(n, dimnum):
    return (n,) * dimnum


def _run(server, reqs, timeout=300):
    outs = {}
    for rid, (params, kw) in reqs.items():
        shape = _shape(params.xDim, params.dimnum)
        outs[rid] = np.zeros(shape, np.complex128)
        server.submit(rid, params, psi_out=outs[rid], **kw)
    server.drain(timeout)
    res = {r["id"]: r for r in server.poll()}
    assert set(res) == set(reqs)
    for r in res.values():
        assert r["status"] == 0, r
    return res, outs


def _mix():
    return {
        1: (_p(128, omega=0.5, corotating=1), dict(seed=11, random_phase=True, max_imag_steps=600, energy_tol=1e-9,
                                                   check_every=100)),
        2: (_p(256), dict(seed=12, max_imag_steps=400, energy_tol=1e-12)),
        3: (_p(64, dimnum=3, g=50.0), dict(seed=13, max_imag_steps=200, real_steps=300, quench_factor=1.5,
                                          check_every=50)),
        4: (_p(192, prec=32), dict(seed=14, max_imag_steps=500, energy_tol=1e-7)),
        5: (_p(512, omega=0.8, corotating=1, xMax=12.0, yMax=12.0), dict(seed=15, random_phase=True,
                                                                         max_imag_steps=300)),
        6: (_p(32, dimnum=3, prec=32), dict(seed=16, max_imag_steps=100, real_steps=200, quench_factor=0.5)),
    }


def test_server_result_same_alone_and_in_mix():
    mix = _mix()
    alone = {}
    for rid, req in mix.items():
        sv = G.Server(budget_bytes=8e9, max_slots=1)
        alone[rid] = _run(sv, {rid: req})
        sv.close()
    sv = G.Server(budget_bytes=64e9, max_slots=16)
    res, outs = _run(sv, mix)
    sv.close()
    for rid in mix:
        ra, oa = alone[rid][0][rid], alone[rid][1][rid]
        rm, om = res[rid], outs[rid]
        assert ra["imag_steps"] == rm["imag_steps"] and ra["real_steps"] == rm["real_steps"], rid
        assert ra["energy"] == rm["energy"] and ra["mu"] == rm["mu"], rid
        assert np.array_equal(oa, om), rid
        assert np.isfinite(rm["energy"]) and np.all(np.isfinite(om))
    # the mix overlapped on the device: requests of different size classes run on
    # different streams, and on the device clock (timing events around each
    # request's GPU work) such a pair was in flight at the same time
    def cls(params):
        n = params.xDim * (params.yDim if params.dimnum > 1 else 1) * (params.zDim if params.dimnum > 2 else 1)
        b = n * (8 if params.precision == 32 else 16)
        return 0 if b >= 64 << 20 else 1 if b >= 8 << 20 else 2 if b >= 1 << 20 else 3
    for r in res.values():
        assert 0.0 <= r["t_dev_start"] < r["t_dev_done"], r
    pairs = [(a, b) for a in mix for b in mix if a < b and cls(mix[a][0]) != cls(mix[b][0])]
    assert any(max(res[a]["t_dev_start"], res[b]["t_dev_start"]) < min(res[a]["t_dev_done"], res[b]["t_dev_done"])
               for a, b in pairs)


def test_server_matches_context_path():
    """A server ground state equals the same chunks driven through gpue_evolve
    and gpue_energy on a context (bit-exact)."""
    params = _p(128, g=300.0)
    sv = G.Server(budget_bytes=4e9, max_slots=4)
    psi0 = np.zeros((128, 128), np.complex128)
    sv.submit(8, params, seed=3, max_imag_steps=0, psi_out=psi0)   # the seeded, normalised psi0 only
    sv.drain(60)
    assert [r["status"] for r in sv.poll()] == [0]
    res, outs = _run(sv, {7: (params, dict(seed=3, max_imag_steps=300, energy_tol=0.0, check_every=100))})
    sv.close()
    assert res[7]["imag_steps"] == 300
    ctx = G.Context(params=params)
    ctx.upload(psi0)
    for k in range(3):
        ctx.evolve(100, 0, first_step=100 * k)
        e, mu = ctx.energy()
    psi = ctx.download()
    ctx.close()
    assert e == res[7]["energy"] and mu == res[7]["mu"]
    assert np.array_equal(psi, outs[7])


def test_server_budget_serialises():
    """With room for one request at a time, admissions follow completions."""
    params = _p(256)
    one = 2 * 16 * 256 * 256 + 4 * 1024 * 1024
    sv = G.Server(budget_bytes=1.5 * one, max_slots=8)
    reqs = {k: (params, dict(seed=20 + k, max_imag_steps=200)) for k in range(4)}
    res, _ = _run(sv, reqs)
    sv.close()
    order = sorted(res.values(), key=lambda r: r["t_start"])
    for a, b in zip(order, order[1:]):
        assert b["t_start"] >= a["t_done"] - 1e-3


def test_server_rejects_oversized():
    sv = G.Server(budget_bytes=1e6, max_slots=2)
    with pytest.raises(G.GpueError):
        sv.submit(1, _p(256))
    sv.close()


def test_server_drain_timeout_status():
    """drain reports GPUE_ETIMEOUT while requests are still running; they finish
    normally afterwards."""
    sv = G.Server(budget_bytes=4e9, max_slots=2)
    sv.submit(1, _p(256, g=300.0), seed=1, max_imag_steps=20000, energy_tol=0.0, check_every=1000)
    with pytest.raises(G.GpueError) as ei:
        sv.drain(0.0)
    assert ei.value.status == G.GPUE_ETIMEOUT
    sv.drain(120)
    res = sv.poll()
    sv.close()
    assert [r["status"] for r in res] == [0] and res[0]["imag_steps"] == 20000
