"""Serving engine: continuous batching of mixed jobs.  CPU tests cover the
stream generator, footprint arithmetic, slot buckets and the shutdown of the
worker threads (with a host-only solver); GPU tests check that every job's
result in the mix equals the same job run alone (time integrations within
1e-12 relative: chunk boundaries move the point where the psi suffix sums are
recomputed; Newton results identical), that jobs arriving while others run
join the running batch, and the same at the batched sweep's sizes with
Newton and continuation jobs at two of its resolutions."""
import threading
import time

import numpy as np
import pytest

from spectraldoublediffusiveconvection_amd import serving


def test_poisson_stream_mix_and_sizes():
    case = dict(N_fm=16, N_r=10, d=2.0, Pr=10.0, Tau=1.0, Ra=6780.0, Ra_s=0.0,
                X=np.zeros(3 * 16 * 9), symmetric=True)
    jobs = serving.poisson_stream(200.0, 2.0, seed=1, newton_case=case, sizes=[(12, 16), (20, 64)])
    kinds = [j.kind for j in jobs]
    assert 300 < len(jobs) < 500
    frac = {k: kinds.count(k) / len(kinds) for k in ("time", "newton", "continuation")}
    assert abs(frac["time"] - 0.6) < 0.1 and abs(frac["newton"] - 0.3) < 0.1
    assert all(np.diff([j.arrival for j in jobs]) > 0)
    cont = [j for j in jobs if j.kind == "continuation"]
    assert cont and all(len(j.ra_path) == 19 and j.arc_points == 0 for j in cont)
    arc = serving.poisson_stream(200.0, 2.0, seed=1, newton_case=case, sizes=[(12, 16)],
                                 arclength=True)
    assert all(j.arc_points == 19 and j.ra_path is None for j in arc if j.kind == "continuation")
    assert serving.class_bytes(512, 40, 64, "mixed") > serving.class_bytes(64, 20, 64, "mixed")


def _solve_alone(j, batch=8):
    """The engine's algorithm on one Newton or continuation job alone: Newton
    solves as a batch of one on a plan reserved like the engine's contexts,
    arclength steps with the single-instance Krylov object."""
    import torch

    from spectraldoublediffusiveconvection_amd import device as dev, newton, operators as ops
    plan = dev.Plan(ops.build(j.N_fm, j.N_r, j.d, 1.0, j.Pr, j.Tau), symmetric=j.symmetric,
                    max_batch=batch)
    nb = newton.NewtonBatch(plan, restart=j.restart, max_batch=batch)
    kr = newton.Krylov(plan, restart=j.restart)
    opts = dict(tol_newton=j.tol_newton, tol_gmres=j.tol_gmres)
    X = torch.as_tensor(np.asarray(j.X0, dtype=np.float64), device="cuda").reshape(1, -1).clone()
    if j.kind == "newton":
        info = nb.solve_(X, j.Ra, j.Ra_s, **opts)[0]
        out = (X[0].cpu().numpy(), info)
    else:
        points, infos = [], []
        ok = True
        for ra in [j.Ra] + list(j.ra_path or []):
            info = nb.solve_(X, ra, j.Ra_s, **opts)[0]
            infos.append(info)
            points.append((ra, X[0].cpu().numpy()))
            if not info["converged"]:
                ok = False
                break
        if ok and j.arc_points:
            Y = torch.cat((X[0], torch.tensor([j.Ra], dtype=torch.float64, device="cuda")))
            Yd, ds = torch.zeros_like(Y), j.ds
            for _ in range(j.arc_points):
                ds, info = kr.arclength_step_(Y, Yd, j.Ra_s, j.sign, ds, **opts)
                infos.append(info)
                if info["status"] != 0:
                    break
                y = Y.cpu().numpy()
                points.append((float(y[-1]), y[:-1]))
        out = (points, infos)
    nb.close()
    kr.close()
    plan.close()
    return out


def _same_as_alone(j):
    res, info = _solve_alone(j)
    assert info == j.info, (j.id, info, j.info)
    if j.kind == "newton":
        assert np.array_equal(res, j.result), j.id
    else:
        assert len(res) == len(j.result), j.id
        for (ra, Xa), (rb, Xb) in zip(res, j.result):
            assert ra == rb and np.array_equal(Xa, Xb), j.id


@pytest.mark.gpu
def test_jobs_same_alone_and_in_the_mix():
    import torch

    from spectraldoublediffusiveconvection_amd import device as dev, operators as ops
    rng = np.random.default_rng(4)
    jobs = []
    for i in range(10):
        nr_, k_ = [(12, 16), (20, 64)][i % 2]
        x = rng.standard_normal(3 * k_ * (nr_ - 1))
        jobs.append(serving.Job("time", k_, nr_, float(rng.uniform(2500, 3000)), 500.0,
                                1e-3 * x / np.linalg.norm(x), nsteps=int(rng.integers(20, 60)),
                                arrival=float(i)))   # virtual clock: one arrival per iteration
    # Newton jobs on a small steady-convection case (wide gap), started near steady state
    Kw, Nrw, d = 16, 10, 2.0
    tsp = dev.Plan.build(Kw, Nrw, d, 0.075, 10.0, 1.0, symmetric=True)
    x0 = rng.random(3 * Kw * (Nrw - 1))
    X = tsp.step(torch.as_tensor(1e-3 * x0 / np.linalg.norm(x0), device="cuda"), 6780.0, 0.0,
                 nsteps=20000)[0].cpu().numpy()
    for i in range(3):
        jobs.append(serving.Job("newton", Kw, Nrw, 6780.0 * (1 + 0.01 * i), 0.0, X, d=d, Pr=10.0,
                                Tau=1.0, symmetric=True, restart=60, arrival=2.0 + i))
    jobs.append(serving.Job("continuation", Kw, Nrw, 6780.0, 0.0, X, d=d, Pr=10.0, Tau=1.0,
                            symmetric=True, restart=60, ra_path=[6800.0, 6820.0], arrival=3.0))
    jobs.append(serving.Job("continuation", Kw, Nrw, 6780.0, 0.0, X, d=d, Pr=10.0, Tau=1.0,
                            symmetric=True, restart=60, arc_points=3, ds=10.0, arrival=3.5))
    eng = serving.Engine(precision="fp64", capacity=4, chunk=7, newton_workers=2)
    done = eng.run(jobs, realtime=False)
    eng.shutdown()
    assert len(done) == len(jobs)
    st = eng.stats()
    assert st["latency_p99_ms"] >= st["latency_p50_ms"] > 0
    # jobs were admitted while earlier ones were still running
    t_jobs = [j for j in jobs if j.kind == "time"]
    assert any(a.t_admit < b.t_done and b.t_admit < a.t_done
               for a in t_jobs for b in t_jobs if a is not b and a.arrival > b.arrival)
    for j in t_jobs:
        plan = dev.Plan(ops.build(j.N_fm, j.N_r, j.d, j.dt, j.Pr, j.Tau), precision="fp64",
                        max_batch=4)
        alone = plan.step(torch.as_tensor(j.X0, device="cuda"), j.Ra, j.Ra_s,
                          nsteps=j.nsteps)[0].cpu().numpy()
        assert np.linalg.norm(j.result - alone) / np.linalg.norm(alone) < 1e-12, j.id
    for j in jobs:
        if j.kind == "time":
            continue
        assert not (isinstance(j.info, dict) and "exception" in j.info), j.info
        _same_as_alone(j)
        if j.kind == "continuation" and j.arc_points:
            ras = [ra for ra, _ in j.result]
            assert len(ras) == 1 + j.arc_points and all(np.diff(ras) > 0), ras
    first = [j for j in jobs if j.kind == "newton"][0]
    assert first.info["converged"], first.info   # from its own steady state


@pytest.mark.gpu
def test_budget_admits_one_class_at_a_time():
    """With an HBM budget that holds one size class, time jobs of two classes
    still all finish: a class is released when it is idle and the other needs
    the memory, at most one class is resident at any time, and every result
    equals the same job run alone."""
    import torch

    from spectraldoublediffusiveconvection_amd import device as dev, operators as ops
    sizes = [(12, 16), (20, 64)]
    cap = 4
    budget = max(serving.class_bytes(k, n, cap, "fp64") for n, k in sizes) + 1024
    rng = np.random.default_rng(7)
    jobs = []
    for i in range(6):
        nr_, k_ = sizes[(i // 2) % 2]
        x = rng.standard_normal(3 * k_ * (nr_ - 1))
        jobs.append(serving.Job("time", k_, nr_, float(rng.uniform(2500, 3000)), 500.0,
                                1e-3 * x / np.linalg.norm(x), nsteps=int(rng.integers(10, 30)),
                                arrival=float(i)))
    eng = serving.Engine(precision="fp64", capacity=cap, chunk=5, newton_workers=1,
                         hbm_budget=budget)
    resident = []
    orig = eng._class_for

    def tracked(job):
        c = orig(job)
        resident.append(len(eng.classes))
        return c

    eng._class_for = tracked
    done = eng.run(jobs, realtime=False)
    eng.shutdown()
    assert len(done) == len(jobs)
    assert resident and max(resident) == 1
    for j in jobs:
        plan = dev.Plan(ops.build(j.N_fm, j.N_r, j.d, j.dt, j.Pr, j.Tau), precision="fp64",
                        max_batch=cap)
        alone = plan.step(torch.as_tensor(j.X0, device="cuda"), j.Ra, j.Ra_s,
                          nsteps=j.nsteps)[0].cpu().numpy()
        assert np.linalg.norm(j.result - alone) / np.linalg.norm(alone) < 1e-12, j.id


def test_slot_buckets():
    assert serving.slot_buckets(4) == [1, 4]
    assert serving.slot_buckets(16) == [1, 4, 16]
    assert serving.slot_buckets(64) == [1, 4, 16, 32, 64]
    assert serving.slot_buckets(256) == [1, 4, 16, 32, 64, 128, 192, 256]
    assert serving.slot_buckets(100) == [1, 4, 16, 32, 64, 100]


def test_shutdown_cancels_queued_jobs_and_joins_workers():
    """A worker inside a (bounded) solve finishes it; the jobs still queued
    are cancelled; shutdown returns only when every worker thread has ended.
    Host-only: the solver is a stand-in that waits on an event."""
    running, release = threading.Event(), threading.Event()
    calls = []

    def solve(worker, jobs):
        calls.extend(j.id for j in jobs)
        running.set()
        release.wait(10.0)
        return [("solved", {"converged": True})] * len(jobs)

    eng = serving.Engine(newton_workers=1, solve=solve)
    jobs = [serving.Job("continuation", 8, 8, 1.0, 0.0, np.zeros(3), arrival=0.0)
            for _ in range(5)]
    eng.run(jobs, realtime=True, max_seconds=0.3)          # returns with the jobs unfinished
    assert running.wait(5.0)
    assert all(not w.daemon for w in eng.workers)
    threading.Timer(0.2, release.set).start()
    t0 = time.perf_counter()
    eng.shutdown()
    assert time.perf_counter() - t0 < 8.0
    assert all(not w.is_alive() for w in eng.workers)
    assert calls == [jobs[0].id]
    assert [j.result for j in eng.done] == ["solved"]
    assert len(eng.cancelled) == 4
    assert all(j.cancelled and j.info == {"cancelled": True} for j in eng.cancelled)
    assert eng.stats()["cancelled"] == 4


@pytest.mark.gpu
def test_sweep_sizes_with_newton_and_continuation_at_two_resolutions():
    """Time jobs drawn from the batched sweep's twelve size classes (the
    40/512 headline class among them) next to Newton and pseudo-arclength
    continuation jobs at two of the sweep's resolutions (N_r, N_fm) = (20, 64)
    and (30, 128), started from the thin-gap convecting state of
    Run_Tests.py:170-196 (d = 0.353, Ra_T ~ 2360, Ra_S = 0, Pr = 1, Tau = 1,
    symmetric).  Every result equals the same job run alone: time
    integrations within 1e-12, Newton and continuation results bit for bit."""
    import torch

    import convecting as C
    from spectraldoublediffusiveconvection_amd import device as dev, operators as ops
    rng = np.random.default_rng(11)
    sizes = [(r, k) for r in (20, 30, 40) for k in (64, 128, 256, 512)]
    jobs = []
    for i, (nr_, k_) in enumerate(sizes):
        x = rng.standard_normal(3 * k_ * (nr_ - 1))
        jobs.append(serving.Job("time", k_, nr_, float(rng.uniform(2500, 3000)), 500.0,
                                1e-3 * x / np.linalg.norm(x), nsteps=(60, 100)[i % 2],
                                arrival=0.5 * i))
    c = C.THIN_GAP
    for j, (Nr, K) in enumerate([(20, 64), (30, 128)]):
        X = C.steady_on_device(K, Nr)      # steady at this resolution
        common = dict(d=c["d"], Pr=c["Pr"], Tau=c["Tau"], symmetric=True, restart=60)
        # close to onset: Ra_T steps of 0.02% at most (Newton needs the
        # reference's limit of 5 iterations once Ra_T moves by 0.04%)
        Xp = X * (1.0 + 2e-3 * np.cos(np.arange(X.size)))
        for q in range(2):   # two Newton jobs of the class arrive together: one batch
            jobs.append(serving.Job("newton", K, Nr, c["Ra"] * (1 + 1e-4 * q), c["Ra_s"], Xp,
                                    arrival=1.0 + j, **common))
        jobs.append(serving.Job("continuation", K, Nr, c["Ra"], c["Ra_s"], X, arc_points=2,
                                ds=10.0, arrival=2.0 + j, **common))
        jobs.append(serving.Job("continuation", K, Nr, c["Ra"], c["Ra_s"], X,
                                ra_path=[c["Ra"] + 0.4], arrival=3.0 + j, **common))
    eng = serving.Engine(precision="mixed", capacity=16, chunk=20, newton_workers=2)
    try:
        done = eng.run(jobs, realtime=False)
    finally:
        eng.shutdown()
    assert len(done) == len(jobs)
    for j in jobs:
        assert not (isinstance(j.info, dict) and "exception" in j.info), j.info
        if j.kind == "time":
            plan = dev.Plan(ops.build(j.N_fm, j.N_r, j.d, j.dt, j.Pr, j.Tau), precision="mixed",
                            max_batch=16)
            alone = plan.step(torch.as_tensor(j.X0, device="cuda"), j.Ra, j.Ra_s,
                              nsteps=j.nsteps)[0].cpu().numpy()
            assert np.linalg.norm(j.result - alone) / np.linalg.norm(alone) < 1e-12, j.id
            plan.close()
            continue
        _same_as_alone(j)
        if j.kind == "newton":
            assert j.info["converged"], j.info
        else:
            assert all(i.get("converged", i.get("status") == 0) for i in j.info), j.info
            assert len(j.result) == 1 + j.arc_points + len(j.ra_path or [])
