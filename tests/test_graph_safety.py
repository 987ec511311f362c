"""HIP graphs and the plan workspace they hold (no counterpart in the
reference, whose time loop is plain Python).

* A graph captures the plan's workspace pointers, so growing the workspace
  (sddc_plan_reserve) or destroying the plan while a graph is alive is
  refused with SDDC_ERR_INVALID instead of leaving the graph on freed memory.
* libsddc holds one device-state lock over a whole capture and over every
  call that allocates, frees, copies synchronously or synchronises the
  device.  Another thread creating, growing and destroying plans and Krylov
  objects (hipMalloc, hipMemcpy, hipFree, hipDeviceSynchronize) while graphs
  are captured must leave every capture valid, and every replay equal to
  the eager steps.  (Kernel launches and waits of other threads are the
  caller's to keep out of a capture: HIP invalidates a capture when another
  thread works on the device meanwhile; the serving engine's Gate does it.)
"""
import ctypes
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from spectraldoublediffusiveconvection_amd import _lib  # noqa: E402
from spectraldoublediffusiveconvection_amd import device as dev  # noqa: E402
from spectraldoublediffusiveconvection_amd import newton  # noqa: E402

CASE = dict(d=0.35, dt=5e-3, Pr=1.0, Tau=1.0 / 15.0)


def _states(B, K, Nr, seed):
    X = np.random.default_rng(seed).standard_normal((B, 3 * K * (Nr - 1)))
    return torch.as_tensor(1e-3 * X / np.linalg.norm(X, axis=1, keepdims=True), device="cuda")


def test_reserve_and_destroy_refused_while_a_graph_holds_the_workspace():
    K, Nr = 32, 12
    plan = dev.Plan.build(K, Nr, **CASE, max_batch=4)
    X = _states(4, K, Nr, 1)
    g = plan.graph(X, 3000.0, 500.0, nsteps=2)
    L = _lib.lib()
    assert L.sddc_plan_reserve(plan._h, 4) == _lib.SDDC_OK        # no growth: nothing freed
    assert L.sddc_plan_reserve(plan._h, 8) == _lib.SDDC_ERR_INVALID
    assert b"graph" in L.sddc_last_error()
    with pytest.raises(ValueError, match="graph"):
        plan.reserve(8)
    assert plan.max_batch == 4
    assert L.sddc_plan_destroy(plan._h) == _lib.SDDC_ERR_INVALID
    g.launch()
    torch.cuda.synchronize()
    g.close()
    assert L.sddc_plan_reserve(plan._h, 8) == _lib.SDDC_OK
    plan.close()


def test_capture_while_another_thread_allocates_and_frees():
    K, Nr, B = 64, 20, 16
    main_plan = dev.Plan.build(K, Nr, **CASE, max_batch=B)
    X0 = _states(B, K, Nr, 2)
    Ra = torch.full((B,), 3000.0, dtype=torch.float64, device="cuda")
    Ras = torch.full((B,), 500.0, dtype=torch.float64, device="cuda")
    # the worker's loop makes only libsddc calls that take the lock
    wK, wNr = 16, 8
    stop, started = threading.Event(), threading.Event()
    counts, errors = {"plans": 0}, []

    def worker():
        try:
            while not stop.is_set():
                p = dev.Plan.build(wK, wNr, **CASE, max_batch=1)    # hipMalloc, hipMemcpy
                p.reserve(2)                                        # hipDeviceSynchronize, hipFree
                kr = newton.Krylov(p, restart=8)                    # hipMalloc, hipHostMalloc
                kr.close()                                          # hipFree
                p.close()                                           # hipFree
                counts["plans"] += 1
                started.set()
        except Exception as e:   # reported by the main thread
            errors.append(e)
            started.set()

    t = threading.Thread(target=worker)
    t.start()
    try:
        started.wait(timeout=60)
        for i in range(6):
            X = X0.clone()
            g = main_plan.graph(X, Ra, Ras, nsteps=3)
            g.launch()
            g.launch()
            torch.cuda.synchronize()
            # each call (and each replay) starts with the suffix sums of its input
            want = main_plan.step(main_plan.step(X0.clone(), Ra, Ras, nsteps=3), Ra, Ras, nsteps=3)
            torch.cuda.synchronize()
            assert torch.equal(X, want), "replay %d differs from the eager steps" % i
            g.close()
    finally:
        stop.set()
        t.join()
    assert not errors, errors
    assert counts["plans"] >= 1
    print("worker created and destroyed %d plans during the captures" % counts["plans"])
