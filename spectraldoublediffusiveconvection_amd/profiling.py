"""Per-kernel timing with HIP events and algorithmic-byte rooflines.

Each stage of the step is launched on torch's current stream between two
events; the algorithmic bytes of a stage are the data it must move through
HBM per launch (operators counted once per launch, shared by the batch).
"""
from __future__ import annotations

import torch

STAGES = {0: "k_suffix", 1: "k_spectral", 2: "k_nonlinear", 3: "k_rhs", 4: "k_solve"}
HBM_PEAK_GBS = 8000.0


def stage_bytes(plan, B, stage):
    K, nr = plan.N_fm, plan.nr
    N = K * nr
    rs = 4 if plan.precision_name == "mixed" else 8
    if stage == 0:
        return B * 2 * N * 8                                   # psi in, suffix sums out
    if stage == 1:
        return B * (4 * N * 8 + 3 * N * 8 + 7 * N * rs)        # X, P in; L, slabs out
    if stage == 2:
        return B * 11 * N * rs                                 # 7 slab rows in, 4 rows out
    if stage == 3:
        return B * (4 * N * rs + 3 * N * 8 + 3 * N * 8)        # F, L in; G out
    # solves: G in, X and P out; the block inverses of the three fields and
    # D2 Linv_j of the A4 chains once per launch
    return B * 7 * N * 8 + 4 * K * nr * nr * 8


def time_stage(plan, X, Ra, Ra_s, stage, reps=20):
    """Average duration (ms) of one stage, HIP events on the current stream.
    Ra, Ra_s should already be device tensors (stage_table converts them), so
    that the timed region holds the stage's kernels and nothing else."""
    for _ in range(2):
        plan.run_stage(X, Ra, Ra_s, stage)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        plan.run_stage(X, Ra, Ra_s, stage)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def stage_table(plan, X, Ra, Ra_s, reps=20):
    """{stage name: (avg_ms, bytes)} for stages 1-4 of one batched step."""
    B = X.shape[0]
    Ra, Ra_s = plan._params(Ra, B), plan._params(Ra_s, B)          # device-resident once
    plan.step(X, Ra, Ra_s, nsteps=1, out=torch.empty_like(X))   # fills the workspace
    out = {}
    for st in (1, 2, 3, 4):
        out[STAGES[st]] = (time_stage(plan, X, Ra, Ra_s, st, reps), stage_bytes(plan, B, st))
    return out


def roofline(plan, X, Ra, Ra_s, reps=20):
    """Roofline entry of the dominant (longest) kernel of the step."""
    tab = stage_table(plan, X, Ra, Ra_s, reps)
    name = max(tab, key=lambda k: tab[k][0])
    ms, nbytes = tab[name]
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": name,
            "avg_us": ms * 1e3, "bytes_per_launch": nbytes,
            "stages_us": {k: v[0] * 1e3 for k, v in tab.items()}}


def copy_probe(nbytes=256 << 20):
    """One device-to-device copy of nbytes: the known-size reference that
    scales rocprofv3's FETCH_SIZE and WRITE_SIZE counters to HBM bytes."""
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print("copied %d bytes" % nbytes)


if __name__ == "__main__":
    copy_probe()
