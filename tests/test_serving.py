"""Serving engine: continuous batching of mixed jobs.  CPU tests cover the
stream generator and footprint arithmetic; GPU tests check that every job's
result in the mix equals the same job run alone (time integrations within
1e-12 relative: chunk boundaries move the point where the psi suffix sums are
recomputed; Newton results identical), and that jobs arriving while others
run join the running batch."""
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
    from spectraldoublediffusiveconvection_amd import newton
    for j in jobs:
        if j.kind == "time":
            continue
        assert not (isinstance(j.info, dict) and "exception" in j.info), j.info
        plan = dev.Plan(ops.build(j.N_fm, j.N_r, j.d, 1.0, j.Pr, j.Tau), symmetric=True)
        kr = newton.Krylov(plan, restart=j.restart)
        # the same solve alone: identical iterates, so identical state and info
        if j.kind == "newton":
            Xa, info = kr.newton(torch.as_tensor(j.X0, device="cuda"), j.Ra, j.Ra_s)
            assert info == j.info, (info, j.info)
            assert np.array_equal(Xa.cpu().numpy(), j.result)
        elif j.arc_points == 0:
            Xa = torch.as_tensor(j.X0, device="cuda")
            for (ra, Xm), ji in zip(j.result, j.info):
                Xa, info = kr.newton(Xa, ra, j.Ra_s)
                assert info == ji, (info, ji)
                assert np.array_equal(Xa.cpu().numpy(), Xm)
        else:
            # Newton at Ra, then arclength steps: the same bits alone, and Ra_T
            # increases along the branch
            assert len(j.result) == 1 + j.arc_points
            Xa, info = kr.newton(torch.as_tensor(j.X0, device="cuda"), j.Ra, j.Ra_s)
            assert info == j.info[0] and np.array_equal(Xa.cpu().numpy(), j.result[0][1])
            Y = torch.cat((Xa, torch.tensor([j.Ra], dtype=torch.float64, device="cuda")))
            Yd, ds = torch.zeros_like(Y), j.ds
            for (ra, Xm), ji in zip(j.result[1:], j.info[1:]):
                Y, Yd, ds, info = kr.arclength_step(Y, Yd, j.Ra_s, j.sign, ds)
                assert info == ji and info["status"] == 0, (info, ji)
                assert np.array_equal(Y.cpu().numpy()[:-1], Xm) and Y[-1].item() == ra
            ras = [ra for ra, _ in j.result]
            assert all(np.diff(ras) > 0), ras
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
