"""Device Newton-Krylov (sddc_gmres, sddc_newton) against dense linear algebra
and the oracle's Newton solve.  GMRES: the Jacobian of a tiny problem is
assembled column by column from the oracle's linearised step and solved with
numpy; the device solve must match it.  Newton: the reference's wide-gap
steady-convection case (Run_Tests.py:140-235) from a time-stepped state."""
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
