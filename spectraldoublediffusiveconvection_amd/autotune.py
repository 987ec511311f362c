"""Autotuner of the plans' launch shapes.

For one size class and batch size it times HIP-graph-replayed steps under each
candidate launch shape (rows per chain segment of the solves, instances per
solve workgroup, sub-batches on forked streams) and keeps the fastest.  The
graphs hold GRAPH_STEPS steps, as the benchmark's and the sweep's do: the
sub-batch streams start together at every graph launch, so the length of the
graph changes how much they overlap.  Results are stored in a JSON table
keyed by size class, precision and batch bucket; ``TuningTable.lookup``
serves them to batching.Sweep and the serving engine.

    python -m spectraldoublediffusiveconvection_amd.autotune [--out table.json]

tunes the sweep's twelve size classes (N_r in {20,30,40} x N_fm in
{64,128,256,512}) for a range of batch sizes.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(PKG, "tuning_mi355x.json")
BUCKETS = (1, 4, 16, 64, 256)
GRAPH_STEPS = 40   # steps per timed graph


def bucket(B):
    """Smallest tuned batch bucket holding B instances."""
    for b in BUCKETS:
        if B <= b:
            return b
    return BUCKETS[-1]


def candidates(N_fm, B=1):
    """Launch shapes tried for a class with N_fm modes (N_fm/2 rows per chain)
    at batch B: quad-lane solves with 1, 2 or 4 instances per workgroup, the
    mixed widths 3, 5 and 6 (2 per A4, 4 per NAB2 workgroup; 5 with the
    two-exchange A4 chains, 6 with three of them per workgroup) and 9, 10
    (one instance per A4 workgroup with deeper prefetch), from B = 16 on the
    16-instance MFMA solve, and from B = 64 on batches split in 2 or 4 parts
    on forked streams of the graph."""
    n = N_fm // 2
    segs = {n, max(8, 4 * math.ceil(2 * math.sqrt(n) / 4))}
    if B < 64:   # short segments only pay off while the device is not full
        segs.add(max(8, 4 * math.ceil(math.sqrt(n) / 4)))
    segs = sorted(segs)
    nbs = (1, 2, 3, 4, 5, 6, 9, 10, 16) if B >= 16 else (1, 2, 3, 4, 5, 6, 9, 10)
    subs = (1, 2, 4) if B >= 64 else (1,)
    return [{"chain_segment": m, "solve_instances": nb, "sub_batches": sb}
            for m in segs for nb in nbs for sb in subs]


def shape_key(t):
    """Timing key of a launch shape."""
    k = "seg%d_nb%d_sb%d" % (t["chain_segment"], t["solve_instances"], t.get("sub_batches", 1))
    if t.get("spectral_instances", 0):
        k += "_sp%d" % t["spectral_instances"]
    return k


def spectral_per_group(N_fm, B, spectral_instances):
    """Instances one spectral workgroup steps through at batch B: the
    library's automatic choice for 0 (from the (mode tile, instance) pairs),
    and never more than B."""
    if spectral_instances <= 0:
        tiles = (N_fm + 15) // 16
        return max(1, min(4, tiles * B // 512))
    return min(spectral_instances, B)


def refine_class(table, N_fm, N_r, B, widths=(), precision="mixed", steps=GRAPH_STEPS,
                 margin=0.05, spectral=(), rounds=3):
    """Time the table's shape for one class and batch again beside variants
    of it -- the given solve widths at the candidate segment lengths, and the
    given spectral-stage instances per workgroup -- and keep the fastest; a
    new shape must win by `margin`.  The shapes are timed in `rounds`
    interleaved rounds and each keeps its best.  Repeated plans of one shape
    at batches of 1-16 differ by up to 15% (each plan gets new allocations),
    hence the 5% margin; spectral variants that run the same launch as the
    table's shape are not timed.  Every timing of the entry comes from this
    run.  Returns (tuning, {shape: µs})."""
    cur = table.lookup((N_fm, N_r), B, precision)
    if cur is None:
        return tune_class(N_fm, N_r, B, precision, steps=steps)
    shapes = {shape_key(cur): cur}
    for m in sorted({c["chain_segment"] for c in candidates(N_fm, B)}):
        for nb in widths:
            t = dict(cur, chain_segment=m, solve_instances=nb)
            shapes.setdefault(shape_key(t), t)
    eff = spectral_per_group(N_fm, B, cur.get("spectral_instances", 0))
    for sp in spectral:
        if spectral_per_group(N_fm, B, sp) == eff:
            continue
        t = dict(cur, spectral_instances=sp)
        shapes.setdefault(shape_key(t), t)
    times = {k: float("inf") for k in shapes}
    for _ in range(max(1, rounds)):
        for k, t in shapes.items():
            times[k] = min(times[k], time_shape(N_fm, N_r, B, precision, t, steps=steps))
    ck = shape_key(cur)
    wins = [k for k in shapes if times[k] < times[ck] * (1.0 - margin)]
    best_k = min(wins, key=times.get) if wins else ck
    return shapes[best_k], {k: round(v, 2) for k, v in times.items()}


def _key(N_fm, N_r, precision, B):
    return "%d,%d,%s,%d" % (N_fm, N_r, precision, bucket(B))


class TuningTable:
    def __init__(self, path=DEFAULT_TABLE):
        self.path = path
        self.entries = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.entries = json.load(f).get("entries", {})

    def lookup(self, key, B, precision):
        """Tuning for a size-class key (N_fm, N_r, ...) at batch B, or None."""
        e = self.entries.get(_key(key[0], key[1], precision, B))
        return dict(e["tuning"]) if e else None

    def put(self, N_fm, N_r, precision, B, tuning, timings):
        self.entries[_key(N_fm, N_r, precision, B)] = {"tuning": tuning, "timings_us": timings}

    def save(self, path=None):
        with open(path or self.path, "w") as f:
            json.dump({"device": "MI355X (gfx950)", "entries": self.entries}, f, indent=1,
                      sort_keys=True)


_OPS = {}


def _ops_for(N_fm, N_r, d, dt):
    from . import operators as ops
    key = (N_fm, N_r, d, dt)
    if key not in _OPS:
        _OPS.clear()
        _OPS[key] = ops.build(N_fm, N_r, d, dt, 1.0, 1.0 / 15.0)
    return _OPS[key]


def time_shape(N_fm, N_r, B, precision, tuning, steps=GRAPH_STEPS, reps=3, d=0.35, dt=5e-3,
               min_seconds=0.03):
    """Best µs per batched step under one launch shape (graph replay): the
    best of `reps` launches, and of as many more as fill `min_seconds`, so
    that the short graphs of small batches are timed over enough launches."""
    import torch

    from . import device as dev
    plan = dev.Plan(_ops_for(N_fm, N_r, d, dt), precision=precision, max_batch=B, tuning=tuning)
    n3 = 3 * N_fm * (N_r - 1)
    x = np.random.default_rng(0).standard_normal((B, n3))
    X = torch.as_tensor(1e-3 * x / np.linalg.norm(x, axis=1, keepdims=True), device="cuda")
    g = plan.graph(X, np.linspace(2500.0, 3000.0, B), 500.0, nsteps=steps)
    g.launch()
    torch.cuda.synchronize()
    best, n, spent = float("inf"), 0, 0.0
    while n < reps or spent < min_seconds:
        t0 = time.perf_counter()
        g.launch()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        best = min(best, el / steps * 1e6)
        n, spent = n + 1, spent + el
    g.close()
    plan.close()
    return best


def tune_class(N_fm, N_r, B, precision="mixed", steps=GRAPH_STEPS):
    """Best launch shape for one class and batch size: (tuning, {shape: µs})."""
    timings = {}
    best, best_t = None, float("inf")
    for t in candidates(N_fm, B):
        us = time_shape(N_fm, N_r, B, precision, t, steps=steps)
        timings[shape_key(t)] = round(us, 2)
        if us < best_t:
            best, best_t = t, us
    return best, timings


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--N_r", type=int, nargs="*", default=[20, 30, 40])
    ap.add_argument("--N_fm", type=int, nargs="*", default=[64, 128, 256, 512])
    ap.add_argument("--batches", type=int, nargs="*", default=list(BUCKETS))
    ap.add_argument("--steps", type=int, default=GRAPH_STEPS, help="steps per timed graph")
    ap.add_argument("--refine", type=int, nargs="*", default=None,
                    help="only compare the table's shapes with these solve widths")
    ap.add_argument("--refine-spectral", type=int, nargs="*", default=None,
                    help="only compare the table's shapes with these spectral instances")
    args = ap.parse_args()
    table = TuningTable(args.out)
    for N_r in args.N_r:
        for N_fm in args.N_fm:
            for B in args.batches:
                if args.refine is not None or args.refine_spectral is not None:
                    t, tm = refine_class(table, N_fm, N_r, B, args.refine or (), args.precision,
                                         steps=args.steps, spectral=args.refine_spectral or ())
                else:
                    t, tm = tune_class(N_fm, N_r, B, args.precision, steps=args.steps)
                table.put(N_fm, N_r, args.precision, B, t, tm)
                print("N_r=%d N_fm=%d B=%d -> %s  %s" % (N_r, N_fm, B, t, tm), flush=True)
                table.save()
    table.save()


if __name__ == "__main__":
    main()
