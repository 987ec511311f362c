"""Benchmark of the pseudospectral double-diffusive time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]

A "step" is one batched step of B instances of the size class N_r=40,
N_fm=512 (d=0.35, Pr=1, Tau=1/15, dt=5e-3) in mixed precision (float32
transforms, products and Chebyshev GEMMs; float64 linear terms and solves),
replayed from a HIP graph.  Ra_T is spread over [2500, 3000] and Ra_s is 500
(at Ra_T = 9000 the conductive state grows like exp(~20 t) and overflows
within a few hundred steps; over [2500, 3000] the run stays bounded).  States are
seeded normal vectors scaled to ||X||_2 = 1e-3, resident in HBM before the
timed region.  ``value`` is the aggregate instance-steps per second.

Beside it the line reports the plain float64 single-instance step (the
baseline the metric is quoted against), the roofline of the dominant kernel
and the CPU baseline (the oracle's NumPy step on one host core, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = dict(N_r=40, N_fm=512, d=0.35, Pr=1.0, Tau=1.0 / 15.0, dt=5e-3, Ra_s=500.0,
                Ra_lo=2500.0, Ra_hi=3000.0)
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--graph-steps", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=100.0)
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    return ap.parse_args()


def seeded_states(B, n3, seed=1234):
    X = np.random.default_rng(seed).standard_normal((B, n3))
    return 1e-3 * X / np.linalg.norm(X, axis=1, keepdims=True)


def measured_traffic(kernel, B):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE runs (scaled by a known-size copy), when they were
    taken on this workload; None otherwise."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("workload") != "N_r=40 N_fm=512 B=%d mixed" % B:
        return None
    return rec["traffic_bytes"]


def sweep_c4(tuning_table, n=256, graph_steps=40, launches=2):
    """Mixed-resolution sweep: n instances with N_r in {20,30,40} and N_fm in
    {64,128,256,512} (seeded), one HIP graph over all size classes."""
    import torch

    from spectraldoublediffusiveconvection_amd import batching
    w = WORKLOAD
    insts = batching.random_sweep(n, seed=2024, Ra=(w["Ra_lo"], w["Ra_hi"]))
    sw = batching.Sweep(insts, precision="mixed", tuning_table=tuning_table)
    g = sw.graph(nsteps=graph_steps)
    g.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(launches):
        g.launch()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    finite = all(bool(torch.isfinite(c.X).all().item()) for c in sw.classes.values())
    g.close()
    return {"instances": n, "size_classes": len(sw.classes), "graph_steps": graph_steps,
            "instance_steps_per_s": n * graph_steps * launches / el,
            "ms_per_sweep_step": 1e3 * el / (graph_steps * launches), "states_finite": finite}


_WIDE_GAP = {}


def serving_run(rate, seconds, tuning_table):
    """Poisson stream of mixed jobs through the serving engine: 60% time
    integrations (500 steps, sizes as in the sweep), 30% Newton solves and 10%
    20-point pseudo-arclength continuations in Ra_T of steady wide-gap
    convection (d=2, Pr=10, N_r=16, N_fm=32, started near the steady state)."""
    import torch

    from spectraldoublediffusiveconvection_amd import device as dev
    from spectraldoublediffusiveconvection_amd import serving
    Kw, Nrw, d = 32, 16, 2.0
    if "X" not in _WIDE_GAP:   # steady-state start of the Newton and continuation jobs
        tsp = dev.Plan.build(Kw, Nrw, d, 0.075, 10.0, 1.0, symmetric=True)
        x0 = np.random.default_rng(0).random(3 * Kw * (Nrw - 1))
        _WIDE_GAP["X"] = tsp.step(torch.as_tensor(1e-3 * x0 / np.linalg.norm(x0), device="cuda"),
                                  6780.0, 0.0, nsteps=26667)[0].cpu().numpy()
    X = _WIDE_GAP["X"]
    case = dict(N_fm=Kw, N_r=Nrw, d=d, Pr=10.0, Tau=1.0, Ra=6780.0, Ra_s=0.0, X=X,
                symmetric=True, restart=100, dRa=10.0)
    w = WORKLOAD
    jobs = serving.poisson_stream(rate, seconds, seed=7, newton_case=case,
                                  ra=(w["Ra_lo"], w["Ra_hi"]), arclength=True)
    eng = serving.Engine(precision="mixed", capacity=64, chunk=25, newton_workers=3,
                         tuning_table=tuning_table)
    try:
        eng.run(jobs, realtime=True, max_seconds=seconds * 10 + 30)
    finally:   # the worker threads end before the interpreter, also on an error
        eng.shutdown()
    st = eng.stats()
    st.update({"rate_jobs_per_s": rate, "arrival_window_s": seconds, "submitted": len(jobs),
               "mix": "60% time (500 steps), 30% newton, "
                      "10% pseudo-arclength continuation (20 points)"})
    return st


def cpu_baseline(seconds):
    """The oracle's float64 NumPy step (reference algorithm restated) on one
    host core, over a bounded sample of the same workload."""
    from threadpoolctl import threadpool_limits

    from oracle import sddc_oracle as O
    w = WORKLOAD
    o = O.build_operators(w["N_fm"], w["N_r"], w["d"], w["dt"], w["Pr"], w["Tau"])
    X = seeded_states(1, 3 * w["N_fm"] * (w["N_r"] - 1))[0]
    with threadpool_limits(1):
        X = O.step(X, o, w["Ra_hi"], w["Ra_s"])   # first step untimed (reference timer, Main.py:325)
        n, t0 = 0, time.perf_counter()
        while True:
            X = O.step(X, o, w["Ra_hi"], w["Ra_s"])
            n += 1
            el = time.perf_counter() - t0
            if el >= seconds and n >= 5:
                break
    return {"value": n / el, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": "%d float64 steps of one N_r=40, N_fm=512 instance (NumPy restatement "
                      "of the reference step, one thread), %.1f s" % (n, el)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (instances are independent; no multi-GPU path); --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch

    from spectraldoublediffusiveconvection_amd import autotune
    from spectraldoublediffusiveconvection_amd import device as dev
    from spectraldoublediffusiveconvection_amd import operators as ops
    from spectraldoublediffusiveconvection_amd import profiling

    w = WORKLOAD
    B = args.batch
    op = ops.build(w["N_fm"], w["N_r"], w["d"], w["dt"], w["Pr"], w["Tau"])
    n3 = 3 * w["N_fm"] * (w["N_r"] - 1)
    Ra = np.linspace(w["Ra_lo"], w["Ra_hi"], B)
    table = autotune.TuningTable()
    tuning = table.lookup((w["N_fm"], w["N_r"]), B, "mixed")

    # ---- batched mixed-precision step, graph-replayed
    plan = dev.Plan(op, precision="mixed", max_batch=B, tuning=tuning)
    X = torch.as_tensor(seeded_states(B, n3), device="cuda")
    gsteps = max(1, min(args.graph_steps, args.steps))
    graph = plan.graph(X, Ra, w["Ra_s"], nsteps=gsteps)
    for _ in range(max(1, args.warmup // gsteps)):
        graph.launch()
    torch.cuda.synchronize()
    nl = max(1, args.steps // gsteps)
    steps = nl * gsteps
    t0 = time.perf_counter()
    for _ in range(nl):
        graph.launch()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    finite = bool(torch.isfinite(X).all().item())
    value = B * steps / el

    # ---- per-kernel timing of the dominant kernel (HIP events on its stream)
    roof = profiling.roofline(plan, X, Ra, w["Ra_s"], reps=20)
    roof["traffic"] = measured_traffic(roof["kernel"], B)

    # ---- plain float64 single-instance step (eager launches)
    p64 = dev.Plan(op)
    x1 = torch.as_tensor(seeded_states(1, n3, seed=99), device="cuda")
    out = torch.empty_like(x1)
    p64.step(x1, w["Ra_hi"], w["Ra_s"], nsteps=5, out=out)
    torch.cuda.synchronize()
    n64 = 200
    t1 = time.perf_counter()
    p64.step(x1, w["Ra_hi"], w["Ra_s"], nsteps=n64, out=out)
    torch.cuda.synchronize()
    el64 = time.perf_counter() - t1
    plain = n64 / el64

    # ---- mixed method against float64: kinetic energy and Nusselt numbers
    # every 100 steps over 10^4 steps of one instance
    p32 = dev.Plan(op, precision="mixed")
    x64 = torch.as_tensor(seeded_states(1, n3, seed=5), device="cuda")
    x32 = x64.clone()
    worst = np.zeros(3)
    for _ in range(100):
        x64 = p64.step(x64, w["Ra_hi"], w["Ra_s"], nsteps=100)
        x32 = p32.step(x32, w["Ra_hi"], w["Ra_s"], nsteps=100)
        d64 = p64.diagnostics(x64)[0, 1:4].cpu().numpy()
        d32 = p64.diagnostics(x32)[0, 1:4].cpu().numpy()
        worst = np.maximum(worst, np.abs(d32 / d64 - 1.0))
    mixed_err = {"steps": 10000, "every": 100, "ke_rel_max": float(worst[0]),
                 "nu_t_minus_1_rel_max": float(worst[1]), "nu_s_minus_1_rel_max": float(worst[2])}

    line = {
        "metric": "aggregate CN/AB2 steps/s at Nr=40,Ntheta=512 vs plain FP64 GPU; p50/p99 job latency",
        "value": value,
        "unit": "instance-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "N_r=40 N_fm=512 d=0.35 Pr=1 Tau=1/15 dt=5e-3 Ra_s=500 "
                               "Ra_T in [2500,3000]; batched mixed precision, HIP graph",
                   "batch": B, "graph_steps": gsteps, "states_finite": finite,
                   "launch_shape": plan.tuning()},
        "plain_fp64_gpu": {"steps_per_s": plain, "ms_per_step": 1e3 * el64 / n64},
        "speedup_vs_plain_fp64": value / plain,
        "mixed_vs_fp64": mixed_err,
        "roofline": roof,
    }
    if not args.no_sweep:
        line["sweep_mixed_sizes"] = sweep_c4(table)
    if not args.no_serving:
        line["serving"] = serving_run(args.serve_rate, args.serve_seconds, table)
        line["serving_rates"] = []
        for rate in (25.0, 200.0):
            st = serving_run(rate, 3.0, table)
            line["serving_rates"].append({k: st[k] for k in (
                "rate_jobs_per_s", "submitted", "jobs", "jobs_per_s", "latency_p50_ms",
                "latency_p99_ms", "instance_steps_per_s")})
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
