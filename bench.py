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
baseline the metric is quoted against), the same batch in float64, the
roofline of the dominant kernel (at the graph's sub-batch shape and on the
full batch), the mixed-vs-FP64 error in convecting flow, the mixed-size
sweep, the serving legs and the CPU baseline (the oracle's NumPy step: one
host core, and single-thread processes on every core of the affinity set,
bounded samples).
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
    ap.add_argument("--cpu-agg-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the mixed-vs-float64 run in convecting flow (profiler runs)")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
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
            rec = json.load(f).get(kernel, {}).get(str(B))
    except (OSError, ValueError, AttributeError):
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


HEADLINE_KEY = (512, 40, 0.35, 5e-3, 1.0, 1.0 / 15.0, False)
_KEEP = ("rate_jobs_per_s", "submitted", "jobs", "jobs_per_s", "latency_p50_ms",
         "latency_p99_ms", "instance_steps_per_s")


def serving_headline(tuning_table, rates=(10.0, 25.0, 100.0, 200.0), seconds=3.0,
                     saturation_jobs=1024, capacity=256):
    """The serving engine on a stream of N_r=40, N_fm=512 time integrations
    (500 steps each, mixed precision, Ra_T in [2500, 3000], Ra_s in {300, 500},
    seeded states of norm 1e-3): Poisson arrivals at each rate for `seconds`,
    then `saturation_jobs` jobs that all arrive at once (the engine's
    throughput with every slot busy).  Graphs are captured before the streams."""
    from spectraldoublediffusiveconvection_amd import serving
    eng = serving.Engine(precision="mixed", capacity=capacity, chunk=50, newton_workers=0,
                         tuning_table=tuning_table)
    try:
        eng.prepare(keys=[HEADLINE_KEY])
        out = {"capacity": capacity, "chunk": 50, "job_steps": 500,
               "launch_shape": eng.classes[HEADLINE_KEY].plan.tuning(), "rates": []}
        for rate in rates:
            jobs = serving.poisson_stream(rate, seconds, seed=int(rate), mix=(1.0, 0.0, 0.0),
                                          sizes=[(40, 512)])
            st = _run_engine(eng, jobs, seconds * 10 + 30)
            st.update({"rate_jobs_per_s": rate, "submitted": len(jobs)})
            out["rates"].append({k: st[k] for k in _KEEP})
        rng = np.random.default_rng(3)
        n3 = 3 * 512 * 39
        jobs = []
        for _ in range(saturation_jobs):
            x = rng.standard_normal(n3)
            jobs.append(serving.Job("time", 512, 40, float(rng.uniform(2500.0, 3000.0)),
                                    float(rng.choice((300.0, 500.0))),
                                    1e-3 * x / np.linalg.norm(x), nsteps=500, arrival=0.0))
        st = _run_engine(eng, jobs, 120)
        out["saturation"] = {"jobs": st["jobs"], "elapsed_s": st["elapsed_s"],
                             "instance_steps_per_s": st["instance_steps_per_s"],
                             "jobs_per_s": st["jobs_per_s"],
                             "latency_p50_ms": st["latency_p50_ms"],
                             "latency_p99_ms": st["latency_p99_ms"]}
    finally:
        eng.shutdown()
    return out


def _run_engine(eng, jobs, max_seconds):
    """One stream through an engine that stays up between streams."""
    eng.done, eng.cancelled = [], []
    before = eng.instance_steps
    eng.run(jobs, realtime=True, max_seconds=max_seconds)
    st = eng.stats()
    st["instance_steps_per_s"] = (eng.instance_steps - before) / eng.elapsed
    return st


def newton_cases(dev, torch):
    """Steady thin-gap convection (d=0.353, Ra_T=2360, Ra_S=0, Pr=1, Tau=1,
    equatorially symmetric) at two of the sweep's resolutions, the starting
    states of the serving stream's Newton and continuation jobs.  The case is
    close to onset, where the steady amplitude moves fast with Ra_T: a job's
    Ra_T is within 0.02% of 2360 and its start the steady state perturbed by
    0.1% (Newton at dt = 1 needs the reference's limit of 5 iterations once
    Ra_T moves by 0.04% from a 0.5% perturbed start)."""
    c = THIN_GAP
    cases = []
    for K, Nr in ((64, 20), (128, 30)):
        X, _ = convecting_state(dev, torch, K, Nr)
        cases.append(dict(N_fm=K, N_r=Nr, d=c["d"], Pr=c["Pr"], Tau=c["Tau"], Ra=c["Ra"],
                          Ra_s=c["Ra_s"], X=X[0].cpu().numpy(), symmetric=True, restart=100,
                          dRa=10.0, ra_spread=2e-4, perturb=1e-3))
    return cases


def serving_mixed(tuning_table, cases, rates=(25.0, 100.0, 200.0), seconds=3.0):
    """The C5 stream: Poisson arrivals of 60% time integrations (500 steps,
    sizes drawn from the sweep's twelve classes), 30% Newton solves and 10%
    20-point pseudo-arclength continuations in Ra_T, the last two at
    (N_r, N_fm) = (20, 64) and (30, 128) from steady thin-gap convection."""
    from spectraldoublediffusiveconvection_amd import serving
    w = WORKLOAD
    keys = [(k, r, w["d"], w["dt"], w["Pr"], w["Tau"], False)
            for r in (20, 30, 40) for k in (64, 128, 256, 512)]
    eng = serving.Engine(precision="mixed", capacity=64, chunk=50, newton_workers=8,
                         tuning_table=tuning_table)
    res = []
    try:
        probe = serving.poisson_stream(1000.0, 1.0, seed=1, newton_case=cases, arclength=True)
        eng.prepare(keys=keys, newton_jobs=[j for j in probe if j.kind != "time"])
        for rate in rates:
            jobs = serving.poisson_stream(rate, seconds, seed=7, newton_case=cases,
                                          ra=(w["Ra_lo"], w["Ra_hi"]), arclength=True)
            st = _run_engine(eng, jobs, seconds * 10 + 30)
            newton = [j for j in jobs if j.kind == "newton" and j.t_done is not None]
            cont = [j for j in jobs if j.kind == "continuation" and j.t_done is not None]
            st.update({"rate_jobs_per_s": rate, "submitted": len(jobs),
                       "newton_converged": sum(bool(isinstance(j.info, dict)
                                                    and j.info.get("converged")) for j in newton),
                       "newton_done": len(newton),
                       "continuation_points": sum(len(j.result or []) for j in cont),
                       "continuation_done": len(cont),
                       "mix": "60% time (500 steps), 30% newton, 10% pseudo-arclength "
                              "continuation (20 points)"})
            res.append(st)
    finally:
        eng.shutdown()
    return res


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_steps(seconds, seed=1234):
    """(steps, seconds) of the oracle's float64 NumPy step on one thread: the
    reference timer's definition, step only, first step untimed (Main.py:285-290,
    325-327)."""
    from threadpoolctl import threadpool_limits

    from oracle import sddc_oracle as O
    w = WORKLOAD
    o = O.build_operators(w["N_fm"], w["N_r"], w["d"], w["dt"], w["Pr"], w["Tau"])
    X = seeded_states(1, 3 * w["N_fm"] * (w["N_r"] - 1), seed=seed)[0]
    with threadpool_limits(1):
        X = O.step(X, o, w["Ra_hi"], w["Ra_s"])
        n, t0 = 0, time.perf_counter()
        while True:
            X = O.step(X, o, w["Ra_hi"], w["Ra_s"])
            n += 1
            el = time.perf_counter() - t0
            if el >= seconds and n >= 5:
                return n, el


def _cpu_worker(args):
    seconds, seed = args
    return _cpu_steps(seconds, seed)


def cpu_baseline(seconds, agg_seconds):
    """The oracle's float64 NumPy step (reference algorithm restated) on the
    host: one instance on one thread, then P single-thread processes (one
    instance each) at once, P = the CPUs this process may run on, at most 16
    (the host share of one GPU).  Runs before the GPU is touched, so the
    worker processes are forked from a process without GPU state."""
    import multiprocessing as mp
    n1, el1 = _cpu_steps(seconds)
    affinity = len(os.sched_getaffinity(0))
    P = max(1, min(affinity, 16))
    with mp.get_context("fork").Pool(P) as pool:
        res = pool.map(_cpu_worker, [(agg_seconds, 1234 + i) for i in range(P)])
    agg = sum(n / el for n, el in res)
    return {"value": agg, "unit": "steps/s", "cores": P, "kind": "port",
            "single_core_steps_per_s": n1 / el1,
            "cpu_model": _cpu_model(), "affinity_cpus": affinity,
            "sample": "aggregate of %d single-thread processes, one N_r=40, N_fm=512 instance "
                      "each, %.0f s each (%d steps in all); single core: %d float64 steps in "
                      "%.1f s (NumPy restatement of the reference step)"
                      % (P, agg_seconds, sum(n for n, _ in res), n1, el1)}


THIN_GAP = dict(d=0.353, Ra=2360.0, Ra_s=0.0, Pr=1.0, Tau=1.0)


def convecting_state(dev, torch, K, Nr, seed=1234, steps=30000):
    """Steady convecting state of the reference's thin-gap validation case
    (Run_Tests.py:170-196: d=0.353, Ra_T=2360, Ra_S=0, Pr=1, Tau=1) at (K, Nr):
    a seeded state of norm 1e-3, equatorially symmetric as the reference runs
    the case, stepped with dt=0.075 until steady (the fixed point of the step
    does not depend on dt)."""
    c = THIN_GAP
    sp = dev.Plan.build(K, Nr, c["d"], 0.075, c["Pr"], c["Tau"], symmetric=True)
    x0 = np.random.default_rng(seed).random(3 * K * (Nr - 1))
    X = torch.as_tensor(1e-3 * x0 / np.linalg.norm(x0), device="cuda")[None]
    X = sp.step(X, c["Ra"], c["Ra_s"], nsteps=steps)
    Y = sp.step(X, c["Ra"], c["Ra_s"], nsteps=100)
    drift = float((Y - X).norm() / X.norm())
    sp.close()
    return Y, drift


def mixed_vs_fp64(dev, torch):
    """Mixed precision against float64 at N_r=40, N_fm=512 in convecting flow:
    kinetic energy and Nu - 1 (T and S at R_1) every 100 steps over 10^4 steps
    of dt=5e-3 (symmetric=False), from the thin-gap steady convecting state."""
    w, c = WORKLOAD, THIN_GAP
    K, Nr, dt = w["N_fm"], w["N_r"], w["dt"]
    X, drift = convecting_state(dev, torch, K, Nr)
    p64 = dev.Plan.build(K, Nr, c["d"], dt, c["Pr"], c["Tau"])
    p32 = dev.Plan.build(K, Nr, c["d"], dt, c["Pr"], c["Tau"], precision="mixed")
    ratio = dt * float(p64.nlin_fx(X).norm() / X.norm())
    d0 = p64.diagnostics(X)[0].cpu().numpy()
    x64, x32 = X.clone(), X.clone()
    worst, worst_state = np.zeros(3), 0.0
    for _ in range(100):
        x64 = p64.step(x64, c["Ra"], c["Ra_s"], nsteps=100)
        x32 = p32.step(x32, c["Ra"], c["Ra_s"], nsteps=100)
        d64 = p64.diagnostics(x64)[0, 1:4].cpu().numpy()
        d32 = p64.diagnostics(x32)[0, 1:4].cpu().numpy()
        worst = np.maximum(worst, np.abs(d32 / d64 - 1.0))
        worst_state = max(worst_state, float((x32 - x64).norm() / x64.norm()))
    return {"regime": "thin-gap steady convection (Run_Tests.py:170-196): d=0.353 Ra_T=2360 "
                      "Ra_S=0 Pr=1 Tau=1, dt=5e-3, symmetric=False",
            "dt_F_over_X": ratio, "kinetic_energy": float(d0[1]), "nu_t_minus_1": float(d0[2]),
            "spinup_drift_per_100_steps": drift,
            "steps": 10000, "every": 100, "ke_rel_max": float(worst[0]),
            "nu_t_minus_1_rel_max": float(worst[1]), "nu_s_minus_1_rel_max": float(worst[2]),
            "state_rel_max": worst_state}


def time_graph(graph, steps, gsteps, warmup, torch):
    for _ in range(max(1, warmup // gsteps)):
        graph.launch()
    torch.cuda.synchronize()
    nl = max(1, steps // gsteps)
    t0 = time.perf_counter()
    for _ in range(nl):
        graph.launch()
    torch.cuda.synchronize()
    return nl * gsteps, time.perf_counter() - t0


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (instances are independent; no multi-GPU path); --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    # the CPU baseline first: its worker processes fork from a process that
    # has not touched the GPU yet
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds, args.cpu_agg_seconds)

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
    gsteps = max(1, min(args.graph_steps, args.steps))

    # ---- batched mixed-precision step, graph-replayed
    plan = dev.Plan(op, precision="mixed", max_batch=B, tuning=tuning)
    X = torch.as_tensor(seeded_states(B, n3), device="cuda")
    graph = plan.graph(X, Ra, w["Ra_s"], nsteps=gsteps)
    steps, el = time_graph(graph, args.steps, gsteps, args.warmup, torch)
    finite = bool(torch.isfinite(X).all().item())
    value = B * steps / el
    graph.close()

    # ---- the same batch in float64 (batching alone, without the precision split)
    plan_f64 = dev.Plan(op, precision="fp64", max_batch=B, tuning=tuning)
    Xd = torch.as_tensor(seeded_states(B, n3), device="cuda")
    gd = plan_f64.graph(Xd, Ra, w["Ra_s"], nsteps=gsteps)
    steps_d, el_d = time_graph(gd, args.steps, gsteps, args.warmup, torch)
    batched_fp64 = {"instance_steps_per_s": B * steps_d / el_d, "ms_per_step": 1e3 * el_d / steps_d,
                    "states_finite": bool(torch.isfinite(Xd).all().item())}
    gd.close()
    plan_f64.close()

    # ---- per-kernel timing of the dominant kernel (HIP events on its stream):
    # at the sub-batch shape the graph runs, and on the whole batch
    parts = max(1, (plan.tuning() or {}).get("sub_batches", 1))
    sub = B // parts if B % parts == 0 else B
    roof = profiling.roofline(plan, X[:sub], Ra[:sub], w["Ra_s"], reps=20)
    roof["traffic"] = measured_traffic(roof["kernel"], sub)
    roof["batch"] = sub
    roof_full = profiling.roofline(plan, X, Ra, w["Ra_s"], reps=20)
    roof_full["traffic"] = measured_traffic(roof_full["kernel"], B)
    roof_full["batch"] = B

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

    batched_fp64["speedup_vs_plain_fp64"] = batched_fp64["instance_steps_per_s"] / plain
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
        "batched_fp64_graph": batched_fp64,
        "mixed_over_batched_fp64": value / batched_fp64["instance_steps_per_s"],
        "mixed_vs_fp64": None if args.no_accuracy else mixed_vs_fp64(dev, torch),
        "roofline": roof,
        "roofline_full_batch": roof_full,
    }
    if not args.no_sweep:
        line["sweep_mixed_sizes"] = sweep_c4(table)
    if not args.no_serving:
        sh = serving_headline(table, seconds=args.serve_seconds)
        sh["saturation"]["speedup_vs_plain_fp64"] = sh["saturation"]["instance_steps_per_s"] / plain
        line["serving_40_512"] = sh
        line["serving_mixed"] = serving_mixed(table, newton_cases(dev, torch),
                                              seconds=args.serve_seconds)
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line))


if __name__ == "__main__":
    main()
