#!/usr/bin/env python3
"""Split-operator step benchmark on one MI355X.

Workload (the configuration the headline metric is quoted on): the 2D rotating
condensate, complex128 psi on 512^2 over [-12,12)^2, V = (x^2+y^2)/2 (+ GPUE's
1/2 |A|^2), g = 1000, Omega = 0.9 applied as A = Omega(-y, x) through the
per-axis gauge operators; by default the full job of 4000 imaginary-time steps
at dt = 1e-3 followed by 10000 real-time steps at dt = 1e-4, from a seeded
random-phase Gaussian.  A step is one split-operator time step.  Inputs are
resident in HBM when the timed region starts.

Also reported (extra): the 3D ground state on 256^3 (V = (x^2+y^2+4z^2)/2,
g = 5000, dt = 5e-4) in FP64 and FP32-bulk/FP64-reductions, the FP32/FP64
speedup and energy error, per-kernel costs (each distinct sweep replayed back
to back from a graph, HIP events on the context stream) and the roofline of
the dominant kernel, single-solve latency of the 256^2 ground state, 64
batched uneven FP32 solves, a Poisson serving stream, and a CPU baseline (the
FP64 oracle on the host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def seeded_gaussian(shape, grids, seed, phase, noise):
    rng = np.random.default_rng(seed)
    mesh = np.meshgrid(*grids[::-1], indexing="ij")
    r2 = sum(m ** 2 for m in mesh)
    psi = np.exp(-r2 / 2) * (1 + noise * rng.uniform(-1, 1, shape))
    if phase:
        psi = psi * np.exp(1j * rng.uniform(0, 2 * np.pi, shape))
    return psi.astype(np.complex128)


CFG2 = dict(dimnum=2, xDim=512, yDim=512, xMax=12.0, yMax=12.0, g=1000.0, gpe=1, corotating=1, omega=0.9,
            gdt=1e-3, dt=1e-4)
CFG3 = dict(dimnum=3, xDim=256, yDim=256, zDim=256, xMax=8.0, yMax=8.0, zMax=8.0, g=5000.0, gpe=1, omegaZ=2.0,
            gdt=5e-4)
CFG1 = dict(dimnum=2, xDim=256, yDim=256, xMax=10.0, yMax=10.0, g=500.0, gpe=1, gdt=1e-3)


def make_ctx(G, prec, cfg, seed=1, phase=False, noise=0.01):
    ctx = G.Context(precision=prec, **cfg)
    grids = [ctx.grid(a) for a in range(ctx.dimnum)]
    psi = seeded_gaussian(ctx.shape, grids, seed, phase, noise)
    dV = 1.0
    for a in range(ctx.dimnum):
        dV *= grids[a][1] - grids[a][0]
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * dV)
    ctx.upload(psi)
    return ctx, psi


def tune(ctx, args, prec=64):
    ctx.set_graph_steps(args.graph_steps)
    spec = args.variants32 if (prec == 32 and args.variants32) else args.variants
    if spec:
        v = [int(a) for a in spec.split(",")]
        ctx.set_variants(v)
        return tuple(v)
    return ctx.autotune()


def run_schedule(ctx, imag, real):
    if imag:
        ctx.evolve(imag, 0)
    if real:
        ctx.evolve(real, 1)


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE
# (each scaled by its reading for a plain 256 MiB device copy: 2.0 and 1.0), per
# kernel instance and workload; the passes and tables are in profiles/.
MEASURED_TRAFFIC = {
    "gpue::col_sweep<double, 512, 5, 2>@512sq": 8787797,
    "gpue::col_sweep<double, 512, 5, 3>@512sq": 8784887,
    "gpue::col_sweep<double, 512, 5, 5>@512sq": 8697118,
    "gpue::col_sweep<double, 512, 5, 6>@512sq": 8936286,
    "gpue::row_sweep<double, 512, 3, 0>@512sq": 8717749,
    "gpue::row_sweep<double, 512, 3, 1>@512sq": 8972530,
    "gpue::row_sweep<double, 512, 3, 2>@512sq": 8793964,
    "gpue::row_sweep<double, 512, 3, 7>@512sq": 8841794,
    "gpue::col_sweep<double, 256, 0, 0>@256cube": 536992027,
    "gpue::col_sweep<double, 256, 0, 1>@256cube": 536993838,
    "gpue::col_sweep<double, 256, 0, 2>@256cube": 537408139,
    "gpue::col_sweep<double, 256, 0, 3>@256cube": 537402176,
    "gpue::col_sweep<double, 256, 0, 5>@256cube": 537285972,
    "gpue::col_sweep<double, 256, 0, 6>@256cube": 537569081,
    "gpue::col_sweep<double, 256, 2, 2>@256cube": 537376379,
    "gpue::col_sweep<float, 256, 0, 0>@256cube": 268548237,
    "gpue::col_sweep<float, 256, 0, 1>@256cube": 268545933,
    "gpue::col_sweep<float, 256, 0, 2>@256cube": 268798608,
    "gpue::row_sweep<double, 256, 1, 0>@256cube": 537095067,
    "gpue::row_sweep<double, 256, 1, 1>@256cube": 537334821,
    "gpue::row_sweep<double, 256, 1, 2>@256cube": 538243034,
    "gpue::row_sweep<double, 256, 1, 65>@256cube": 537835606,
    "gpue::row_sweep<float, 256, 0, 0>@256cube": 268648607,
    "gpue::row_sweep<float, 256, 0, 65>@256cube": 268896874,
}

RM = {"row_fwd": 0, "row_inv": 1, "row_gauge": 2, "row_full": 3, "row_ksum": 4, "row_inv_fwd": 5,
      "row_gauge_fwd": 6, "row_k": 7}
CM = {"fwd": 0, "inv": 1, "k": 2, "gauge": 3, "ksum": 4, "fwdv": 5, "invvg": 6, "gauge_fwd": 7}


def kernel_symbol(name, prec, n_axis, variant):
    """The template instance rocprofv3 reports for a sweep name (row sweeps
    are instantiated per plan and twiddle loading, variants 0..3 with 4, 5
    acting as 2, 3; strided sweeps per full variant). Closing row sweeps
    without the gauge operator after them ("+gx") run the instance compiled
    without it (mode bit 64)."""
    t = "double" if prec == 64 else "float"
    parts = name.split("+")
    base = parts[0]
    if base.startswith("row"):
        mode = RM[base]
        if base in ("row_inv", "row_inv_fwd") and "gx" not in parts[1:]:
            mode |= 64
        return f"gpue::row_sweep<{t}, {n_axis[0]}, {(variant[0] & 1) + (2 if variant[0] > 1 else 0)}, {mode}>"
    axis = 1 if base.startswith("coly") else 2
    return f"gpue::col_sweep<{t}, {n_axis[axis]}, {variant[axis]}, {CM[base.split('_', 1)[1]]}>"


def kernel_table(ctx, prec, variants, steps_by_gstate, reps=50):
    """Per-kernel cost from gpue_sweep_profile: each distinct sweep replayed back
    to back from a graph and timed with HIP events on the context stream."""
    S = 16 if prec == 64 else 8
    bytes_per_launch = 2 * S * ctx.N
    dims = [ctx.params.xDim, ctx.params.yDim, ctx.params.zDim]
    rows = []
    for gstate, nsteps in steps_by_gstate.items():
        if nsteps <= 0:
            continue
        for k in ctx.sweep_profile(gstate, reps):
            rows.append(dict(kernel=k["name"], gstate=gstate, symbol=kernel_symbol(k["name"], prec, dims, variants),
                             us_per_launch=round(k["us"], 3), launches_per_step=k["per_step"],
                             job_ms=round(k["us"] * k["per_step"] * nsteps / 1e3, 3),
                             GBps=round(bytes_per_launch / (k["us"] * 1e-6) / 1e9, 1)))
    rows.sort(key=lambda r: -r["job_ms"])
    return rows, bytes_per_launch


def roofline(rows, bytes_per_launch, workload):
    if not rows:   # --no-profile
        return None
    top = rows[0]
    achieved = bytes_per_launch / (top["us_per_launch"] * 1e-6) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": MEASURED_TRAFFIC.get(f"{top['symbol']}@{workload}"), "kernel": top["kernel"],
            "symbol": top["symbol"], "bytes_per_launch": bytes_per_launch,
            "avg_launch_us": top["us_per_launch"]}


def physical_cores():
    """Physical cores of the host per lscpu (distinct core, socket pairs), or None."""
    import subprocess

    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
    except Exception:
        return None
    pairs = {line.strip() for line in out.splitlines() if line and not line.startswith("#")}
    return len(pairs) or None


def cpu_baseline(blocks=5, block_steps=50):
    """The FP64 oracle (OpenMP) on the host: the headline job's imaginary-time
    and real-time steps timed separately in blocks of block_steps steps (median
    of `blocks` blocks each), with every thread this process may use and with
    one; the job rate is 14000 / (4000 / r_imag + 10000 / r_real)."""
    from oracle import oracle as O

    share = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(int(env), share) if env.isdigit() and int(env) > 0 else share
    orc = O.Oracle(**CFG2)
    grids = [orc.grid(0), orc.grid(1)]
    psi = seeded_gaussian((512, 512), grids, 1, True, 0.0)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)

    def rates(nt):
        O.set_threads(nt)
        orc.set_wfc(psi)
        out = {}
        step = 0
        for gstate in (0, 1):
            per = []
            for _ in range(blocks):
                t0 = time.perf_counter()
                for i in range(block_steps):
                    orc.step(step + i, gstate)
                per.append(block_steps / (time.perf_counter() - t0))
                step += block_steps
            out[gstate] = float(np.median(per))
        job = 14000.0 / (4000.0 / out[0] + 10000.0 / out[1])
        return job, out

    t0 = time.perf_counter()
    job_n, ph_n = rates(threads)
    job_1, ph_1 = rates(1)
    O.set_threads(threads)
    el = time.perf_counter() - t0
    return {"value": round(job_n, 3), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (oracle/gpue_oracle.c, OpenMP, {threads} threads) on the 512^2 gauge job: "
                      f"median of {blocks} blocks of {block_steps} imaginary-time steps ({ph_n[0]:.1f} steps/s) and "
                      f"of {blocks} blocks of {block_steps} real-time steps ({ph_n[1]:.1f} steps/s), combined with the "
                      f"job's 4000/10000 split; {el:.1f} s of CPU timing in total",
            "one_core": {"value": round(job_1, 3), "imag_steps_per_s": round(ph_1[0], 3),
                         "real_steps_per_s": round(ph_1[1], 3)},
            "imag_steps_per_s": round(ph_n[0], 3), "real_steps_per_s": round(ph_n[1], 3),
            "physical_cores_lscpu": physical_cores(), "cpus_allowed": share}


def cpu_other_configs():
    """The FP64 oracle on the 256^2 ground state (2D, g=500) and the 256^3 ground
    state (3D, g=5000), imaginary time, with every thread this process may use
    and with one: steps/s as the median of timed blocks.  256^2: 5 blocks of 50
    steps each; 256^3 (about a second per step on 16 threads): 3 blocks of 2 steps
    on all threads and 1 step on one core, so the default run stays short."""
    from oracle import oracle as O

    share = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(int(env), share) if env.isdigit() and int(env) > 0 else share
    out = {}
    for name, cfg, plan in (("256sq_ground_state", CFG1, {threads: (5, 50), 1: (5, 50)}),
                            ("256cube_ground_state", CFG3, {threads: (3, 2), 1: (1, 1)})):
        orc = O.Oracle(**cfg)
        shape = tuple(int(cfg[k]) for k in ("xDim", "yDim", "zDim")[: cfg["dimnum"]])
        grids = [orc.grid(a) for a in range(cfg["dimnum"])]
        psi = seeded_gaussian(shape, grids, 1, False, 0.01)
        psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * orc.dV)
        res = {}
        for nt, (blocks, steps) in plan.items():
            O.set_threads(nt)
            orc.set_wfc(psi)
            per = []
            for b in range(blocks):
                t0 = time.perf_counter()
                for i in range(steps):
                    orc.step(b * steps + i, 0)
                per.append(steps / (time.perf_counter() - t0))
            res["threads_%d" % nt] = {"steps_per_s": round(float(np.median(per)), 3),
                                      "sample": f"median of {blocks} blocks of {steps} imaginary-time steps"}
        out[name] = res
    O.set_threads(threads)
    return out


def config4_specs(n_inst=64, seed=4):
    """The batched sweep's instances: (sizes, [(problem, psi0 seed)]).  Grids
    are a seeded mix of 128^2-512^2, g ~ U[100,2000], Omega ~ U[0,0.95] through
    the gauge operators, omega_y/omega_x ~ U[0.8,1.2]."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([128, 192, 256, 384, 512], n_inst)
    specs = []
    for k in range(n_inst):
        n = int(sizes[k])
        L = 10.0 if n <= 256 else 12.0
        cfg = dict(dimnum=2, xDim=n, yDim=n, xMax=L, yMax=L, g=float(rng.uniform(100, 2000)), gpe=1,
                   omega=float(rng.uniform(0, 0.95)), corotating=1, omegaX=1.0,
                   omegaY=float(rng.uniform(0.8, 1.2)), gdt=1e-3, dt=1e-4)
        specs.append((cfg, 100 + k))
    return sizes, specs


def batched_sweep(G, n_inst=64, seed=4, max_steps=20000, tol=1e-8, force_mode=None, force_variants=None):
    """64 independent 2D imaginary-time solves on uneven grids, FP32 bulk with
    FP64 norm/energy, advanced together through one batch (one launch per sweep
    per shape, two steps per graph replay) until each instance's energy changes
    by less than tol (relative) over 100 steps (at most max_steps).  psi0 is
    the noisy Gaussian exp(-r^2/2)(1 + 0.01 U[-1,1]), seeded per instance."""
    sizes, specs = config4_specs(n_inst, seed)
    ctxs = []
    for k, (cfg, psi_seed) in enumerate(specs):
        ctx, _ = make_ctx(G, 32, cfg, seed=psi_seed, phase=False, noise=0.01)
        if force_variants is not None:
            ctx.set_variants(list(force_variants))
        ctxs.append(ctx)
    B = G.Batch(ctxs)
    # the batch tunes itself as a whole (a lone instance's choice serves latency;
    # in the batch the groups share the device): variants per shape group and
    # the graph mode, timed over 20 batch steps each, states restored after.
    # Forced choices skip it (its timing replays graphs, which kernel tracers avoid).
    tune_us = None
    if force_mode is None and force_variants is None:
        t_tune = time.perf_counter()
        tune_us, best_mode = B.autotune(0, 20)
        t_tune = time.perf_counter() - t_tune
    else:
        best_mode = 0 if force_mode is None else force_mode
        B.set_graph_mode(best_mode)
        t_tune = 0.0
    tuned = {}
    for k, c in enumerate(ctxs):
        tuned.setdefault(int(sizes[k]), c.variants())
    active = [True] * n_inst
    e_prev = [None] * n_inst
    last_change = [None] * n_inst
    steps = [0] * n_inst
    point_steps = 0.0
    bytes_total = 0.0
    per_step_bytes = [c.step_bytes(0)[0] for c in ctxs]
    t0 = time.perf_counter()
    done_at = [None] * n_inst
    while any(active) and max(steps) < max_steps:
        B.evolve(100, 0)
        energies, _ = B.energy()   # every instance at once (one wait)
        for k, c in enumerate(ctxs):
            if not active[k]:
                continue
            steps[k] += 100
            point_steps += 100 * c.N
            bytes_total += 100 * per_step_bytes[k]
            e = float(energies[k])
            if e_prev[k] is not None:
                last_change[k] = abs(e - e_prev[k]) / abs(e_prev[k])
            if e_prev[k] is not None and abs(e - e_prev[k]) < tol * abs(e_prev[k]):
                active[k] = False
                done_at[k] = time.perf_counter() - t0
                B.set_active(k, False)
            e_prev[k] = e
    el = time.perf_counter() - t0
    B.close()
    for c in ctxs:
        c.close()
    conv = sum(1 for d in done_at if d is not None)
    # convergence against the rotation rate: rotating instances nucleate vortices
    # and relax them slowly in imaginary time
    by_omega = {}
    for lo, hi in ((0.0, 0.3), (0.3, 0.6), (0.6, 0.95)):
        ks = [k for k, (cfg, _) in enumerate(specs) if lo <= cfg["omega"] < hi or (hi == 0.95 and cfg["omega"] >= hi)]
        ch = [last_change[k] for k in ks if last_change[k] is not None]
        by_omega[f"{lo}-{hi}"] = {"instances": len(ks), "converged": sum(1 for k in ks if done_at[k] is not None),
                                  "median_last_rel_change_per_100_steps": float(np.median(ch)) if ch else None}
    return {"instances": n_inst, "sizes": {int(n): int(np.sum(sizes == n)) for n in sorted(set(sizes))},
            "converged": conv, "max_steps": max_steps, "wall_s": round(el, 3),
            "instance_steps_per_s": round(sum(steps) / el, 1),
            "grid_point_steps_per_s": round(point_steps / el, 1),
            "effective_GBps": round(bytes_total / el / 1e9, 1),
            "variants_by_size": {int(k): list(v) for k, v in tuned.items()}, "graph_mode": best_mode,
            "batch_autotune_us_per_batch_step": round(tune_us, 1) if tune_us is not None else None,
            "batch_autotune_s": round(t_tune, 2),
            "steps_to_converge": {"min": int(min(steps)), "max": int(max(steps))},
            "convergence_by_omega": by_omega,
            "last_rel_energy_change_per_100_steps": {
                "median": float(np.median([x for x in last_change if x is not None])),
                "max": float(max(x for x in last_change if x is not None))}}


def serving_stream(G, n_req=48, rate=200.0, seed=5, budget=250e9, slots=8):
    """Poisson arrivals (mean rate req/s) of 2D ground states (128^2-1024^2) and
    3D quench runs (64^3-256^3: imaginary-time preparation, g quenched by
    U[0.5,2], 1000 real-time steps), each in FP64 or FP32-bulk, admitted by the
    HBM budget; reports throughput and p50/p99 latency (submit to result)."""
    rng = np.random.default_rng(seed)
    sv = G.Server(budget_bytes=budget, max_slots=slots)
    reqs = []
    for k in range(n_req):
        prec = 64 if rng.uniform() < 0.5 else 32
        if rng.uniform() < 0.6:
            n = int(rng.choice([128, 192, 256, 384, 512, 1024]))
            L = 10.0 if n <= 256 else 12.0
            p = G.make_params(dimnum=2, xDim=n, yDim=n, xMax=L, yMax=L, g=float(rng.uniform(100, 1000)), gpe=1,
                              gdt=1e-3, dt=1e-4, precision=prec)
            kw = dict(max_imag_steps=2000, energy_tol=1e-10 if prec == 64 else 1e-6, check_every=100)
            kind = "2d_ground"
        else:
            n = int(rng.choice([64, 128, 192, 256]))
            p = G.make_params(dimnum=3, xDim=n, yDim=n, zDim=n, xMax=8.0, yMax=8.0, zMax=8.0,
                              g=float(rng.uniform(100, 2000)), gpe=1, omegaZ=2.0, gdt=5e-4, dt=1e-4,
                              precision=prec)
            kw = dict(max_imag_steps=100, energy_tol=0.0, real_steps=1000,
                      quench_factor=float(rng.uniform(0.5, 2.0)), check_every=100)
            kind = "3d_quench"
        reqs.append((k, p, kw, kind, n, prec))
    gaps = rng.exponential(1.0 / rate, n_req)
    t0 = time.perf_counter()
    t_next = t0
    for (k, p, kw, kind, n, prec), gap in zip(reqs, gaps):
        t_next += gap
        while time.perf_counter() < t_next:
            time.sleep(min(1e-3, max(0.0, t_next - time.perf_counter())))
        sv.submit(k, p, seed=1000 + k, **kw)
    sv.drain(900.0)
    res = sv.poll(4096)
    el = time.perf_counter() - t0
    sv.close()
    lat = np.array([r["t_done"] - r["t_submit"] for r in res]) * 1e3
    bad = [r for r in res if r["status"] != 0]
    kinds = {}
    for (k, p, kw, kind, n, prec) in reqs:
        kinds[kind] = kinds.get(kind, 0) + 1

    def pct(a):
        a = np.asarray(a, dtype=float)
        return {"p50": round(float(np.percentile(a, 50)), 2), "p99": round(float(np.percentile(a, 99)), 2),
                "max": round(float(a.max()), 2)}

    # per request class: latency (submit to result) and device time (first kernel to the
    # end of the last chunk, from the request's own events)
    info = {k: (kind, n, prec) for (k, p, kw, kind, n, prec) in reqs}
    by_class = {}
    for r in res:
        kind, n, prec = info[r["id"]]
        key = f"{kind}_{n}_fp{prec}"
        e = by_class.setdefault(key, {"lat": [], "dev": []})
        e["lat"].append((r["t_done"] - r["t_submit"]) * 1e3)
        e["dev"].append((r["t_dev_done"] - r["t_dev_start"]) * 1e3)
    classes = {k: {"requests": len(v["lat"]), "latency_ms": pct(v["lat"]), "device_ms": pct(v["dev"])}
               for k, v in sorted(by_class.items())}
    return {"requests": n_req, "mean_rate_req_s": rate, "failed": len(bad), "mix": kinds,
            "wall_s": round(el, 3), "throughput_req_s": round(len(res) / el, 2),
            "latency_ms": pct(lat), "by_class": classes,
            "budget_GB": budget / 1e9, "max_slots": slots}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=14000, help="timed steps (first 4000/14000 imaginary, rest real)")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-extra", action="store_true", help="skip the 256^3 / latency extras")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the per-kernel graph replays and the roofline (for runs under a kernel tracer, "
                         "whose graph interception has crashed on the host)")
    ap.add_argument("--only", choices=["512", "256c", "lat", "batch", "serve", "copy"], default=None,
                    help="profiling: run one workload only (copy: a plain 256 MiB device copy, the "
                         "calibration kernel for the HBM byte counters)")
    ap.add_argument("--steps3", type=int, default=200, help="timed 256^3 steps per precision")
    ap.add_argument("--batch-steps", type=int, default=20000, help="batched sweep: most steps per instance")
    ap.add_argument("--batch-mode", type=int, default=None, choices=[0, 1, 2],
                    help="batched sweep: force the graph mode (2: no graphs, for kernel tracers)")
    ap.add_argument("--batch-variants", type=str, default=None, help="batched sweep: force the variants x,y,z")
    ap.add_argument("--graph-steps", type=int, default=32,
                    help="steps per captured graph (library default 8; 32 saves ~1.5 %% of graph launch "
                         "gaps here; 0: every sweep launched individually, e.g. under a tracer)")
    ap.add_argument("--variants32", type=str, default=None,
                    help="kernel variants of the FP32 contexts (default: as --variants)")
    ap.add_argument("--variants", type=str, default=None,
                    help="kernel variants per axis, e.g. 1,1,1 (default: autotune)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every solve runs on a single MI355X "
              "(multi-GPU sharding is a design note only). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import gpue_amd as G

    K, W = args.steps, args.warmup
    imag = int(round(K * 4000 / 14000))
    real = K - imag
    extra = {}

    # ---------------- headline: 512^2 rotating condensate, FP64, gauge ----------------
    if args.only in (None, "512"):
        ctx, psi0 = make_ctx(G, 64, CFG2, phase=True, noise=0.0)
        variants = tune(ctx, args)
        run_schedule(ctx, W // 2, W - W // 2)
        ctx.upload(psi0)
        # the step blocks the timed calls replay are captured here, not inside the timing
        ctx.prepare_evolve(imag, 0)
        ctx.prepare_evolve(real, 1)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.evolve(imag, 0)
        ctx.sync()   # split point for the per-phase rates (one sync in the job)
        t_mid = time.perf_counter()
        ctx.evolve(real, 1)
        ctx.sync()
        t_end = time.perf_counter()
        el = t_end - t0
        sps = K / el
        e_final, mu_final = ctx.energy()
        e_cons, _ = ctx.energy_consistent()   # gauge term of the evolved Hamiltonian (+hbar <A.k>)
        # the same job with every sweep launched on its own (no graphs), for comparison
        ctx.set_graph_steps(0)
        ctx.upload(psi0)
        ctx.sync()
        t_e0 = time.perf_counter()
        run_schedule(ctx, imag, real)
        ctx.sync()
        eager_sps = K / (time.perf_counter() - t_e0)
        ctx.set_graph_steps(args.graph_steps)
        bps_imag, P_imag = ctx.step_bytes(0)
        bps_real, P_real = ctx.step_bytes(1)
        bytes_job = imag * bps_imag + real * bps_real
        rows, bpl = kernel_table(ctx, 64, variants, {0: imag, 1: real}) if not args.no_profile else ([], 0)
        # the boundary takes and returns host buffers: one upload of psi0 and one download
        # of the result around the job give the PCIe-inclusive rate (never the headline)
        ctx.sync()
        t_a = time.perf_counter()
        ctx.upload(psi0)
        ctx.sync()
        t_b = time.perf_counter()
        ctx.download()
        t_c = time.perf_counter()
        pcie = {"upload_ms": round((t_b - t_a) * 1e3, 3), "download_ms": round((t_c - t_b) * 1e3, 3),
                "steps_per_s_incl_transfers": round(K / (el + (t_c - t_a)), 1),
                # GPUE's loop copies psi to the host at the start of every step (src/evolution.cu:430)
                "steps_per_s_if_psi_downloaded_every_step": round(K / (el + K * (t_c - t_b)), 1)}
        ctx.close()
        log(f"512^2 FP64 gauge: {sps:.1f} steps/s ({el / K * 1e6:.2f} us/step; imaginary "
            f"{(t_mid - t0) / max(imag, 1) * 1e6:.2f}, real {(t_end - t_mid) / max(real, 1) * 1e6:.2f}), "
            f"E={e_final:.10g}, variants {variants}")
        for r in rows:
            log(f"   {r['kernel']:24s} g{r['gstate']} {r['us_per_launch']:8.2f} us x {r['launches_per_step']:.1f}/step "
                f"{r['GBps']:8.1f} GB/s  {r['symbol']}")
        roof = roofline(rows, bpl, "512sq")
        extra["512sq_fp64_gauge"] = {
            "steps_per_s": round(sps, 1), "us_per_step": round(el / K * 1e6, 3),
            "us_per_step_imag": round((t_mid - t0) / max(imag, 1) * 1e6, 3),
            "us_per_step_real": round((t_end - t_mid) / max(real, 1) * 1e6, 3),
            "effective_GBps": round(bytes_job / el / 1e9, 1), "sweeps_per_step_imag": P_imag,
            "sweeps_per_step_real": P_real, "energy_final": e_final, "mu_final": mu_final,
            "energy_consistent_final": e_cons, "pcie_inclusive": pcie,
            "graph_steps": args.graph_steps, "eager_steps_per_s": round(eager_sps, 1),
            "variants": list(variants), "kernels": rows}
    else:
        sps, el, roof = float("nan"), float("nan"), None

    # ---------------- 256^3 FP64 vs FP32 ----------------
    if args.only == "copy":
        import torch

        a = torch.ones(1 << 24, dtype=torch.complex128, device="cuda")   # 256 MiB
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy: 5 x 256 MiB device-to-device (268435456 bytes read, 268435456 written per launch)")
        return

    if not args.no_extra and args.only in (None, "256c"):
        K3 = args.steps3
        res3 = {}
        for prec in (64, 32):
            ctx, psi0 = make_ctx(G, prec, CFG3, seed=1)
            variants = tune(ctx, args, prec)
            ctx.evolve(10, 0)
            ctx.upload(psi0)
            ctx.sync()
            t0 = time.perf_counter()
            ctx.evolve(K3, 0)
            ctx.sync()
            el3 = time.perf_counter() - t0
            e3, mu3 = ctx.energy()
            b3, P3 = ctx.step_bytes(0)
            rows3, bpl3 = kernel_table(ctx, prec, variants, {0: K3}, reps=10) if not args.no_profile else ([], 0)
            ctx.close()
            res3[prec] = dict(steps_per_s=round(K3 / el3, 2), us_per_step=round(el3 / K3 * 1e6, 2),
                              effective_GBps=round(b3 * K3 / el3 / 1e9, 1),
                              frac_of_8TBps=round(b3 * K3 / el3 / 1e9 / HBM_PEAK_GBS, 4), sweeps_per_step=P3,
                              energy=e3, mu=mu3, variants=list(variants), kernels=rows3,
                              roofline=roofline(rows3, bpl3, "256cube"))
            log(f"256^3 FP{prec}: {K3 / el3:.1f} steps/s, {b3 * K3 / el3 / 1e9:.0f} GB/s effective, E={e3:.12g}, "
                f"variants {variants}")
            for r in rows3:
                log(f"   {r['kernel']:24s} {r['us_per_launch']:8.2f} us x {r['launches_per_step']:.1f}/step "
                    f"{r['GBps']:8.1f} GB/s  {r['symbol']}")
        extra["256cube_fp64"] = res3[64]
        extra["256cube_fp32"] = res3[32]
        extra["fp32_over_fp64_speedup_256cube"] = round(res3[32]["steps_per_s"] / res3[64]["steps_per_s"], 3)
        extra["fp32_rel_energy_error_256cube"] = abs(res3[32]["energy"] - res3[64]["energy"]) / abs(res3[64]["energy"])
        extra["fp32_rel_mu_error_256cube"] = abs(res3[32]["mu"] - res3[64]["mu"]) / abs(res3[64]["mu"])
        # the same ground state rotating (the optional Omega = 0.6 gauge about z): per-axis x and y
        # gauge operators, 6 sweeps per imaginary step; the same autotuner, imaginary time
        cfg = dict(CFG3, corotating=1, omega=0.6)
        ctx, psi0 = make_ctx(G, 64, cfg, seed=1, phase=True)
        variants = tune(ctx, args, 64)
        ctx.evolve(10, 0)
        ctx.upload(psi0)
        K3g = max(2, K3 // 2)
        ctx.prepare_evolve(K3g, 0)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.evolve(K3g, 0)
        ctx.sync()
        elg = time.perf_counter() - t0
        bg, Pg = ctx.step_bytes(0)
        eg, mug = ctx.energy()
        ctx.close()
        extra["256cube_fp64_gauge"] = dict(steps_per_s=round(K3g / elg, 2), us_per_step=round(elg / K3g * 1e6, 2),
                                          effective_GBps=round(bg * K3g / elg / 1e9, 1), sweeps_per_step=Pg,
                                          energy=eg, mu=mug, variants=list(variants), steps=K3g)
        log(f"256^3 FP64 gauge (Omega=0.6): {K3g / elg:.1f} steps/s, {bg * K3g / elg / 1e9:.0f} GB/s effective, "
            f"{Pg} sweeps/step, variants {variants}")

    if not args.no_extra and args.only in (None, "lat"):
        # single-solve latency of the 256^2 ground state (config 1: 2000 imaginary steps)
        lat = []
        for rep in range(7):
            ctx, _ = make_ctx(G, 64, CFG1, seed=1)
            ctx.sync()
            t0 = time.perf_counter()
            ctx.evolve(2000, 0)
            ctx.energy()
            ctx.sync()
            lat.append(time.perf_counter() - t0)
            ctx.close()
        lat = np.array(lat[1:]) * 1e3
        extra["solve_latency_256sq_2000steps_ms"] = {"p50": round(float(np.percentile(lat, 50)), 3),
                                                     "p99": round(float(np.percentile(lat, 99)), 3),
                                                     "samples": len(lat)}

    # ---------------- batched sweep and serving stream ----------------
    if not args.no_extra and args.only in (None, "batch"):
        extra["batched_sweep_64_uneven_fp32"] = batched_sweep(
            G, max_steps=args.batch_steps, force_mode=args.batch_mode,
            force_variants=[int(v) for v in args.batch_variants.split(",")] if args.batch_variants else None)
        log("batched sweep:", json.dumps(extra["batched_sweep_64_uneven_fp32"]))
    if not args.no_extra and args.only == "serve":
        # a burst far above capacity (throughput, queueing latency), then the
        # same mix offered at about half the measured capacity (service latency)
        extra["serving_stream"] = serving_stream(G)
        log("serving stream:", json.dumps(extra["serving_stream"]))
        light = round(0.5 * extra["serving_stream"]["throughput_req_s"], 1)
        extra["serving_stream_half_load"] = serving_stream(G, rate=light)
        log("serving stream at half load:", json.dumps(extra["serving_stream_half_load"]))
    elif not args.no_extra and args.only is None:
        # the serving stream runs in a process of its own, as a solve service would:
        # in this process, after the sections above have created and destroyed some
        # seventy streams, its small requests ran several times slower
        import subprocess

        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--only", "serve", "--no-cpu"],
                           capture_output=True, text=True, timeout=900)
        sys.stderr.write(r.stderr)
        if r.returncode != 0:
            raise RuntimeError(f"serving stream process failed ({r.returncode})")
        sub = json.loads(r.stdout.strip().splitlines()[-1])["extra"]
        extra["serving_stream"] = sub["serving_stream"]
        extra["serving_stream_half_load"] = sub["serving_stream_half_load"]
        extra["serving_stream"]["process"] = "own"

    cpu = None
    if not args.no_cpu and args.only is None:
        cpu = cpu_baseline()
        log(f"CPU baseline: {cpu['value']} steps/s on {cpu['cores']} threads, {cpu['one_core']['value']} on one")
        if sps == sps:
            extra["speedup_vs_cpu_baseline"] = round(sps / cpu["value"], 1)
        extra["cpu_oracle_other_configs"] = cpu_other_configs()
        log("CPU oracle, other configs:", json.dumps(extra["cpu_oracle_other_configs"]))

    out = {
        "metric": "split-op steps/s (512^2 rotating condensate, FP64, gauge Omega=0.9)",
        "value": round(sps, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(el / K * 1e3, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random-phase Gaussian; fields generated in-kernel)",
        "config": {"workload": "2D rotating condensate 512^2 over [-12,12)^2, g=1000, Omega=0.9 gauge, "
                               f"{imag} imaginary (dt=1e-3) + {real} real (dt=1e-4) steps",
                   "grid": [512, 512], "precision": "fp64", "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
