"""Mixed-resolution sweeps: one HIP graph over several size classes, launch
shapes, and the autotuner's table.  GPU tests check that the graph of a sweep
gives the same bits as stepping each class eagerly, and that every instance
gives the same bits alone (same launch shapes) as in the sweep."""
import json

import numpy as np
import pytest

from spectraldoublediffusiveconvection_amd import autotune


def test_autotune_buckets_and_candidates(tmp_path):
    assert [autotune.bucket(b) for b in (1, 2, 4, 5, 16, 17, 300)] == [1, 4, 4, 16, 16, 64, 256]
    c = autotune.candidates(512)
    segs = sorted({t["chain_segment"] for t in c})
    assert segs == [16, 32, 256] and {t["solve_instances"] for t in c} == {1, 2, 3, 4, 5, 6, 9, 10}
    assert {t["solve_instances"] for t in autotune.candidates(512, 64)} == {1, 2, 3, 4, 5, 6, 9, 10,
                                                                            16}
    t = autotune.TuningTable(str(tmp_path / "t.json"))
    t.put(64, 20, "mixed", 10, {"chain_segment": 8, "solve_instances": 2}, {"x": 1.0})
    t.save()
    again = autotune.TuningTable(str(tmp_path / "t.json"))
    assert again.lookup((64, 20), 12, "mixed") == {"chain_segment": 8, "solve_instances": 2}
    assert again.lookup((64, 20), 40, "mixed") is None
    assert json.load(open(tmp_path / "t.json"))["entries"]


@pytest.mark.gpu
def test_sweep_graph_matches_eager_and_alone():
    import torch

    from spectraldoublediffusiveconvection_amd import batching, device as dev, operators as ops
    insts = batching.random_sweep(9, seed=3, N_r=(12, 20), N_fm=(16, 64))
    tuning = {"chain_segment": 8, "solve_instances": 2, "sub_batches": 2}

    class Fixed:
        def lookup(self, key, B, precision):
            return tuning
    a = batching.Sweep(insts, precision="fp64", tuning_table=Fixed())
    b = batching.Sweep(insts, precision="fp64", tuning_table=Fixed())
    assert len(a.classes) >= 2
    # the same call boundaries on both sides: each call starts by recomputing
    # the psi suffix sums, which later steps take from the previous solve
    g = a.graph(nsteps=5)
    for _ in range(2):
        g.launch()
        b.step(5)
    sa, sb = a.states(), b.states()
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)
    g.close()
    # each instance alone, in its own plan with the same launch shapes
    for inst, x in zip(insts, sa):
        p = dev.Plan(ops.build(inst["N_fm"], inst["N_r"], inst["d"], inst["dt"], inst["Pr"],
                               inst["Tau"]), precision="fp64", tuning=tuning)
        y = torch.as_tensor(inst["X0"], device="cuda")
        for _ in range(2):
            y = p.step(y, inst["Ra"], inst["Ra_s"], nsteps=5)
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy()[0], x)
        assert {k: p.tuning()[k] for k in tuning} == tuning


@pytest.mark.gpu
def test_sweep_graph_with_more_parts_than_queues_matches_eager():
    """More parts than hardware queues: the graph places the parts on lanes
    by their timed cost, each lane running its parts in turn.  Two launches
    of a five-step graph give the same bits as the same steps taken eagerly,
    class by class."""
    import os

    from spectraldoublediffusiveconvection_amd import batching
    insts = batching.random_sweep(30, seed=11, N_r=(12, 20), N_fm=(16, 32, 64))
    tuning = {"chain_segment": 8, "solve_instances": 2, "sub_batches": 1}

    class Fixed:
        def lookup(self, key, B, precision):
            return tuning
    a = batching.Sweep(insts, precision="mixed", tuning_table=Fixed())
    b = batching.Sweep(insts, precision="mixed", tuning_table=Fixed())
    queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) or 4
    assert len(a.classes) > queues, (len(a.classes), queues)
    g = a.graph(nsteps=5)
    for _ in range(2):
        g.launch()
        b.step(5)
    for x, y in zip(a.states(), b.states()):
        assert np.array_equal(x, y)
    g.close()


@pytest.mark.gpu
def test_tuned_plans_agree_with_oracle():
    """Every launch shape the autotuner may pick gives the oracle's step."""
    import torch

    from oracle import sddc_oracle as O
    from spectraldoublediffusiveconvection_amd import device as dev, operators as ops
    K, Nr = 64, 20
    o = O.build_operators(K, Nr, 0.35, 5e-3, 1.0, 1 / 15.0)
    X = 1e-3 * np.random.default_rng(8).standard_normal((3, 3 * K * (Nr - 1)))
    ref = [O.step(x, o, 3000.0, 500.0) for x in X]
    for t in autotune.candidates(K, 16):
        p = dev.Plan(ops.build(K, Nr, 0.35, 5e-3, 1.0, 1 / 15.0), max_batch=3, tuning=t)
        y = p.step(torch.as_tensor(X, device="cuda"), 3000.0, 500.0).cpu().numpy()
        for a, b in zip(y, ref):
            assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-12, t


@pytest.mark.gpu
@pytest.mark.parametrize("sub", [2, 4])
def test_sub_batch_graph_same_bits(sub):
    """A graph whose batch is split into parts on forked streams gives the
    same bits as eager steps of the whole batch (parts of 16-instance
    multiples and a ragged last part)."""
    import torch

    from spectraldoublediffusiveconvection_amd import device as dev, operators as ops
    K, Nr, B = 64, 20, 70
    p = dev.Plan(ops.build(K, Nr, 0.35, 5e-3, 1.0, 1 / 15.0), precision="mixed", max_batch=B,
                 tuning={"sub_batches": sub})
    assert p.tuning()["sub_batches"] == sub
    rng = np.random.default_rng(sub)
    X0 = 1e-3 * rng.standard_normal((B, 3 * K * (Nr - 1)))
    Ra = torch.as_tensor(rng.uniform(2500, 3000, B), device="cuda")
    xg = torch.as_tensor(X0, device="cuda")
    g = p.graph(xg, Ra, 500.0, nsteps=3)
    eager = torch.as_tensor(X0, device="cuda")
    for rep in range(2):
        g.launch()
        eager = p.step(eager, Ra, 500.0, nsteps=3)
        torch.cuda.synchronize()
        assert torch.equal(xg, eager), rep
    g.close()


@pytest.mark.gpu
def test_c4_sweep_256_instances_twelve_classes():
    """The C4 configuration: 256 instances drawn from N_r in {20, 30, 40} x
    N_fm in {64, 128, 256, 512} (twelve size classes), Ra_T in [2500, 3000],
    Ra_S in {300, 500}, mixed precision with the autotuned launch shapes of
    the shipped table, three steps from one graph launch.  For three
    instances of every class: the oracle's three float64 steps within 2e-5
    (mixed) and the same sweep in float64 within 1e-12; and each instance
    alone in its own plan with its class's launch shape gives the same bits
    as in the graph."""
    import torch

    from oracle import sddc_oracle as O
    from spectraldoublediffusiveconvection_amd import batching, device as dev, operators as ops
    insts = batching.random_sweep(256, seed=2024)
    table = autotune.TuningTable()
    sw = batching.Sweep(insts, precision="mixed", tuning_table=table)
    assert len(sw.classes) == 12 and sw.n_instances == 256
    g = sw.graph(nsteps=3)
    g.launch()
    got = sw.states()
    g.close()
    s64 = batching.Sweep(insts, precision="fp64", tuning_table=table)
    g64 = s64.graph(nsteps=3)
    g64.launch()
    got64 = s64.states()
    g64.close()
    picked = {}
    for idx, inst in enumerate(insts):
        picked.setdefault(batching.class_key(inst), []).append(idx)
    worst_mixed, worst64 = 0.0, 0.0
    for key, idxs in picked.items():
        c = sw.classes[key]
        shape = c.plan.tuning()
        tuned = table.lookup(key, c.B, "mixed")
        assert tuned is not None and {k: shape[k] for k in tuned} == tuned, (key, shape, tuned)
        o = O.build_operators(*key)
        for idx in idxs[:3]:
            inst = insts[idx]
            x = inst["X0"]
            for _ in range(3):
                x = O.step(x, o, inst["Ra"], inst["Ra_s"])
            e32 = np.linalg.norm(got[idx] - x) / np.linalg.norm(x)
            e64 = np.linalg.norm(got64[idx] - x) / np.linalg.norm(x)
            worst_mixed, worst64 = max(worst_mixed, e32), max(worst64, e64)
            assert e32 < 2e-5, (key, idx, e32)
            assert e64 < 1e-12, (key, idx, e64)
            alone = dev.Plan(ops.build(*key), precision="mixed", tuning=shape)
            y = alone.step(torch.as_tensor(inst["X0"], device="cuda"), inst["Ra"], inst["Ra_s"],
                           nsteps=3)
            assert np.array_equal(y.cpu().numpy()[0], got[idx]), (key, idx)
            alone.close()
    print("C4 sweep: 12 classes, worst relative error after 3 steps vs oracle: mixed %.2e, "
          "float64 %.2e" % (worst_mixed, worst64))


def test_autotune_refine_keeps_a_shape_unless_beaten_by_the_margin(tmp_path, monkeypatch):
    """refine_class times the table's shape again beside the new widths and
    replaces it only when a new shape is faster by more than the margin; the
    entry's timings all come from that run (no timing is simulated beyond
    the stand-in for the GPU timer)."""
    t = autotune.TuningTable(str(tmp_path / "t.json"))
    cur = {"chain_segment": 16, "solve_instances": 1, "sub_batches": 1}
    t.put(512, 40, "mixed", 1, cur, {"stale": 1.0})
    fake = {1: 100.0, 9: 99.0, 10: 90.0}

    def timer(N_fm, N_r, B, precision, tuning, steps=40):
        return fake[tuning["solve_instances"]] + (0.5 if tuning["chain_segment"] != 16 else 0.0)

    monkeypatch.setattr(autotune, "time_shape", timer)
    best, tm = autotune.refine_class(t, 512, 40, 1, (9, 10))
    assert best == {"chain_segment": 16, "solve_instances": 10, "sub_batches": 1}
    assert tm["seg16_nb1_sb1"] == 100.0 and "stale" not in tm
    fake[10] = 99.0          # within 2% of the current shape: kept
    best, _ = autotune.refine_class(t, 512, 40, 1, (9, 10))
    assert best == cur
