"""Device Newton-Krylov (sddc_gmres, sddc_newton) against dense linear algebra
and the oracle's Newton solve.  GMRES: the Jacobian of a tiny problem is
assembled column by column from the oracle's linearised step and solved with
numpy; the device solve must match it.  Newton: the reference's wide-gap
steady-convection case (Run_Tests.py:140-235) from a time-stepped state."""
import os

import numpy as np
import pytest

from oracle import sddc_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


def test_gmres_matches_dense_solve(torch):
    from spectraldoublediffusiveconvection_amd import device as dev
    from spectraldoublediffusiveconvection_amd import newton
    K, Nr, d, dt = 8, 8, 0.35, 5e-3
    Ra, Ras = 3000.0, 500.0
    o = O.build_operators(K, Nr, d, dt, 1.0, 1 / 15.0)
    n3 = 3 * K * (Nr - 1)
    rng = np.random.default_rng(11)
    X = 1e-2 * rng.standard_normal(n3)
    J = np.empty((n3, n3))
    for c in range(n3):
        e = np.zeros(n3)
        e[c] = 1.0
        J[:, c] = O.step_linearized(e, X, o, Ra, Ras) - e
    b = rng.standard_normal(n3)
    x_ref = np.linalg.solve(J, b)
    plan = dev.Plan.build(K, Nr, d, dt, 1.0, 1 / 15.0)
    kr = newton.Krylov(plan, restart=n3)
    t = lambda a: torch.as_tensor(a, device="cuda")
    x, info = kr.gmres(t(X), t(b), Ra, Ras, rtol=1e-13)
    assert info["converged"]
    assert rel(x.cpu().numpy(), x_ref) < 1e-9
    # restarted (short basis) reaches the same solution
    kr2 = newton.Krylov(plan, restart=20)
    x2, info2 = kr2.gmres(t(X), t(b), Ra, Ras, rtol=1e-12, max_restarts=200)
    assert info2["converged"] and info2["restarts"] > 1
    assert rel(x2.cpu().numpy(), x_ref) < 1e-8


@pytest.mark.parametrize("fd", [False, True])
def test_newton_steady_convection(torch, fd):
    from spectraldoublediffusiveconvection_amd import device as dev
    from spectraldoublediffusiveconvection_amd import newton
    d, Kw, Nr, dt = 2.0, 32, 16, 0.075
    Ra, Ras, Tau, Pr = 6780.0, 0.0, 1.0, 10.0
    n3 = 3 * Kw * (Nr - 1)
    X0 = np.random.default_rng(6).random(n3)
    X0 = 1e-3 * X0 / np.linalg.norm(X0)
    ts = dev.Plan.build(Kw, Nr, d, dt, Pr, Tau, symmetric=True)
    X_ts = ts.step(torch.as_tensor(X0, device="cuda"), Ra, Ras, nsteps=26667)[0]
    # perturb away from the steady state so that Newton has work to do
    X_start = X_ts * (1.0 + 0.02 * torch.cos(torch.arange(n3, device="cuda", dtype=torch.float64)))
    plan = dev.Plan.build(Kw, Nr, d, 1.0, Pr, Tau, symmetric=True)
    kr = newton.Krylov(plan, restart=300)
    Xs, info = kr.newton(X_start, Ra, Ras, tol_newton=1e-8, fd=fd)
    assert info["converged"], info
    assert info["iterations"] >= 2
    Xs = Xs.cpu().numpy()
    o = O.build_operators(Kw, Nr, d, 1.0, Pr, Tau)
    assert rel(O.step(Xs, o, Ra, Ras, symmetric=True), Xs) < 1e-9
    ref = O.newton(X_start.cpu().numpy(), Ra, Ras, Tau, Pr, d, Kw, Nr, True, tol_newton=1e-8)
    assert ref[-1] is True
    assert rel(Xs, ref[0]) < 1e-8


@pytest.mark.parametrize("case", ["wide", "thin"])
@pytest.mark.parametrize("fd", [False, True])
def test_newton_batch_same_bits_alone_and_batched(torch, fd, case):
    """Batched Newton (sddc_newton_batched, device Hessenberg) on five
    instances of one class with different Ra_T and starts, solved together:
    each instance gets the same bits as solved alone (a batch of one), is a
    steady state of the oracle's step, and agrees with the single-instance
    solver (host Hessenberg) to the Newton tolerance.  Cases: the wide-gap
    steady state of Run_Tests.py:140-235 (d = 2, N_r = 16, N_fm = 32) and the
    thin-gap one of Run_Tests.py:170-196 at a sweep resolution (N_r = 20,
    N_fm = 64)."""
    from spectraldoublediffusiveconvection_amd import device as dev
    from spectraldoublediffusiveconvection_amd import newton
    Bn = 5
    if case == "wide":
        z = np.load(os.path.join(os.path.dirname(__file__), "golden", "steady_widegap.npz"))
        K, Nr, d, Pr, Tau, Ra0 = int(z["N_fm"]), int(z["N_r"]), float(z["d"]), float(z["Pr"]), \
            float(z["Tau"]), float(z["Ra"])
        X0 = z["X"]
    else:
        import convecting as C
        K, Nr = 64, 20
        X0, c = C.steady_on_device(K, Nr), C.THIN_GAP
        d, Pr, Tau, Ra0 = c["d"], c["Pr"], c["Tau"], c["Ra"]
    # Ra_T within 0.02% and 0.2% smooth perturbations of the steady state:
    # both cases sit close to a bifurcation, where Newton at dt = 1 (batched or
    # single alike) needs 5 iterations -- the reference's failure -- once Ra_T
    # moves by 0.04% from a 0.5% perturbed start
    idx = np.arange(X0.size)
    starts = np.stack([X0 * (1.0 + 0.002 * np.cos(idx + b)) for b in range(Bn)])
    Ras = Ra0 * (1.0 + 5e-5 * np.arange(Bn))
    plan = dev.Plan.build(K, Nr, d, 1.0, Pr, Tau, symmetric=True, max_batch=8)
    nb = newton.NewtonBatch(plan, restart=60, max_batch=8)
    kr = newton.Krylov(plan, restart=60)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    Xb, infos = nb.solve(t(starts), t(Ras), t(np.zeros(Bn)), tol_newton=1e-9, fd=fd)
    Xb = Xb.cpu().numpy()
    o1 = O.build_operators(K, Nr, d, 1.0, Pr, Tau)
    for b in range(Bn):
        Xa, ia = nb.solve(t(starts[b]), float(Ras[b]), 0.0, tol_newton=1e-9, fd=fd)
        Xs, info_s = kr.newton(t(starts[b]), float(Ras[b]), 0.0, tol_newton=1e-9, fd=fd)
        Xs = Xs.cpu().numpy()
        res_b = rel(O.step(Xb[b], o1, Ras[b], 0.0, symmetric=True), Xb[b])
        res_s = rel(O.step(Xs, o1, Ras[b], 0.0, symmetric=True), Xs)
        print("%s fd=%s instance %d: batched %s residual %.2e | single %s residual %.2e | "
              "difference %.2e" % (case, fd, b, infos[b], res_b, info_s, res_s, rel(Xb[b], Xs)))
        assert ia[0] == infos[b], (b, ia[0], infos[b])
        assert np.array_equal(Xa.cpu().numpy()[0], Xb[b]), b
        assert info_s["converged"], info_s
        assert infos[b]["converged"] and infos[b]["iterations"] >= 2, infos[b]
        assert res_b < 1e-9
        assert rel(Xb[b], Xs) < 1e-8
    nb.close()
    kr.close()
