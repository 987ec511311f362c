"""Long-run parity in a convecting regime at the C3 size (N_r = 40,
N_fm = 512) and the C2 size (N_r = 30, N_fm = 256), plus C2's Newton solve.

Regime: the reference's thin-gap validation case (Run_Tests.py:170-196),
d = 0.353, Ra_T = 2360, Ra_S = 0, Pr = 1, Tau = 1, dt = 5e-3, symmetric =
False, started from its steady convecting state (cached at N_r = 24,
N_fm = 48 by golden/make_convecting.py and moved to each size by
convecting.cheb_resample).  The nonlinear term is a large part of every
update there: ||dt F(X)|| / ||X|| = 0.23 at both sizes (asserted >= 1e-2).

Tolerances (relative):
  * one float64 step, GPU vs oracle from the same state:      1e-12
  * kinetic energy, Nu - 1 (T and S at R_1), ||X|| and four state
    projections, GPU float64 vs the oracle's trajectory, every 100 steps
    over 2000 steps (C3) and 10^4 steps (C2):                   1e-9
  * mixed precision vs float64, every 100 steps over 10^4 steps (C3):
    kinetic energy 1e-6, Nu - 1 1e-5, state 1e-5; mixed vs the oracle's
    trajectory over its 2000 steps: kinetic energy 1e-6
  * C2 Newton (finite-difference J v, GMRES restart 50, tol 1e-10):
    ||Step(X*) - X*|| / ||X*|| <= 1e-9 with the oracle's step at dt = 1
"""
import numpy as np
import pytest

import convecting as C
from oracle import sddc_oracle as O


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def fix():
    return C.load()


def _oracle_ops(K, Nr, dt=C.DT):
    c = C.THIN_GAP
    return O.build_operators(K, Nr, c["d"], dt, c["Pr"], c["Tau"])


# ----------------------------------------------------------------- CPU pins
def test_cheb_resample_exact_on_polynomials():
    K0, Nr0, K1, Nr1 = 6, 12, 10, 17
    o0, o1 = _oracle_ops(K0, Nr0), _oracle_ops(K1, Nr1)

    def field(R, K):
        f = (R - o0.R[0]) * (R - o0.R[-1]) * (R ** 5 - 3.0 * R)   # degree 7, zero at the walls
        return np.outer(1.0 / np.arange(1, K + 1) ** 2, f)

    X = np.concatenate([field(o0.R[1:-1], K0).ravel() * s for s in (1.0, -2.0, 0.5)])
    Y = C.cheb_resample(X, K0, Nr0, K1, Nr1)
    ref = np.zeros((K1, Nr1 - 1))
    ref[:K0] = field(o1.R[1:-1], K0)
    want = np.concatenate([ref.ravel() * s for s in (1.0, -2.0, 0.5)])
    assert np.abs(Y - want).max() < 1e-11 * np.abs(want).max()
    # same resolution: identity
    assert np.array_equal(C.cheb_resample(X, K0, Nr0, K0, Nr0), X)


def test_fixture_is_a_steady_convecting_state(fix):
    K, Nr = C.FIXTURE_SIZE
    c = C.THIN_GAP
    o = _oracle_ops(K, Nr, dt=0.075)
    X = fix["X24"]
    assert rel(O.step(X, o, c["Ra"], c["Ra_s"], symmetric=True), X) < 1e-13
    assert np.linalg.norm(X) > 1.0                    # far from the conduction state
    for tag in C.RUNS:
        r = fix[tag + "_ratio"]
        print(tag, "||dt F(X)|| / ||X|| along the oracle trajectory:", np.round(r, 4))
        assert r.min() >= 1e-2


# ---------------------------------------------------------------- GPU parity
def _gpu():
    torch = pytest.importorskip("torch")
    from spectraldoublediffusiveconvection_amd import device as dev
    return torch, dev


def _plan(dev, K, Nr, precision="fp64", dt=C.DT):
    c = C.THIN_GAP
    return dev.Plan.build(K, Nr, c["d"], dt, c["Pr"], c["Tau"], precision=precision)


def _diag(plan, X):
    """||X||, kinetic energy, Nu_T - 1, Nu_S - 1 at R_1 (the oracle fixture's columns)."""
    return plan.diagnostics(X)[0, :4].cpu().numpy()


def _trajectory(torch, plan, X0, nsteps, keep_every=None):
    c = C.THIN_GAP
    X = torch.as_tensor(X0, device="cuda")[None].clone()
    diag, kept = [_diag(plan, X)], {0: X[0].cpu().numpy()}
    for n in range(100, nsteps + 1, 100):
        X = plan.step(X, c["Ra"], c["Ra_s"], nsteps=100)
        diag.append(_diag(plan, X))
        if keep_every and n % keep_every == 0:
            kept[n] = X[0].cpu().numpy()
    return np.array(diag), X[0].cpu().numpy(), kept


def _compare_to_oracle(diag, X_end, fix, tag, tol):
    gold = fix[tag + "_diag"]
    assert diag.shape == gold.shape
    err = np.abs(diag / gold - 1.0).max(axis=0)
    print(tag, "max relative error every 100 steps (||X||, KE, Nu_T-1, Nu_S-1):", err)
    assert err.max() < tol, err
    proj = np.array([C.projection_vector(X_end.size, j) @ X_end for j in range(4)])
    assert rel(proj, fix[tag + "_proj"]) < tol


@pytest.fixture(scope="module")
def c3(fix):
    torch, dev = _gpu()
    K, Nr, nsteps = C.RUNS["c3"]
    X0 = C.start_state(K, Nr, fix)
    p64 = _plan(dev, K, Nr)
    diag, X_end, kept = _trajectory(torch, p64, X0, nsteps, keep_every=500)
    return dict(X0=X0, p64=p64, diag=diag, X_end=X_end, kept=kept, K=K, Nr=Nr)


@pytest.mark.gpu
def test_c3_one_step_parity_along_the_trajectory(c3):
    """GPU float64 step vs the oracle's from the same states, sampled every 500
    steps of the GPU trajectory; the nonlinear share of each update reported."""
    torch, _ = _gpu()
    c = C.THIN_GAP
    o = _oracle_ops(c3["K"], c3["Nr"])
    for n, S in sorted(c3["kept"].items()):
        got = c3["p64"].step(torch.as_tensor(S, device="cuda"), c["Ra"], c["Ra_s"])[0]
        want = O.step(S, o, c["Ra"], c["Ra_s"])
        e = rel(got.cpu().numpy(), want)
        ratio = C.DT * np.linalg.norm(O.nlin_fx(S, o)) / np.linalg.norm(S)
        print("step %5d: one-step error %.2e, ||dt F|| / ||X|| = %.4f" % (n, e, ratio))
        assert e < 1e-12
        assert ratio >= 1e-2


@pytest.mark.gpu
def test_c3_fp64_2000_steps_vs_oracle_trajectory(c3, fix):
    _compare_to_oracle(c3["diag"], c3["X_end"], fix, "c3", 1e-9)


@pytest.mark.gpu
def test_c3_mixed_vs_fp64_10k_steps(c3, fix):
    """Mixed precision (float32 transforms, products and Chebyshev GEMMs of the
    nonlinear term) against float64: kinetic energy and Nusselt numbers every
    100 steps over 10^4 steps, in the convecting regime."""
    torch, dev = _gpu()
    c = C.THIN_GAP
    p32 = _plan(dev, c3["K"], c3["Nr"], precision="mixed")
    p64 = c3["p64"]
    x64 = torch.as_tensor(c3["X0"], device="cuda")[None].clone()
    x32 = x64.clone()
    worst, worst_state, vs_oracle = np.zeros(3), 0.0, np.zeros(3)
    gold = fix["c3_diag"]
    for i in range(1, 101):
        x64 = p64.step(x64, c["Ra"], c["Ra_s"], nsteps=100)
        x32 = p32.step(x32, c["Ra"], c["Ra_s"], nsteps=100)
        d64, d32 = _diag(p64, x64)[1:], _diag(p64, x32)[1:]
        worst = np.maximum(worst, np.abs(d32 / d64 - 1.0))
        worst_state = max(worst_state, float((x32 - x64).norm() / x64.norm()))
        if i < len(gold):
            vs_oracle = np.maximum(vs_oracle, np.abs(d32 / gold[i, 1:] - 1.0))
    print("mixed vs float64 over 10^4 steps, max relative error: KE %.2e, Nu_T-1 %.2e, "
          "Nu_S-1 %.2e, state %.2e; mixed vs oracle over 2000 steps: KE %.2e, Nu_T-1 %.2e"
          % (*worst, worst_state, vs_oracle[0], vs_oracle[1]))
    assert worst[0] <= 1e-6
    assert worst[1:].max() <= 1e-5
    assert worst_state <= 1e-5
    assert vs_oracle[0] <= 1e-6


@pytest.fixture(scope="module")
def c2(fix):
    torch, dev = _gpu()
    K, Nr, nsteps = C.RUNS["c2"]
    X0 = C.start_state(K, Nr, fix)
    p64 = _plan(dev, K, Nr)
    diag, X_end, _ = _trajectory(torch, p64, X0, nsteps)
    return dict(diag=diag, X_end=X_end, K=K, Nr=Nr)


@pytest.mark.gpu
def test_c2_fp64_10k_steps_vs_oracle_trajectory(c2, fix):
    _compare_to_oracle(c2["diag"], c2["X_end"], fix, "c2", 1e-9)


@pytest.mark.gpu
def test_c2_newton_finite_difference_restart_50(c2):
    """C2's steady solve: matrix-free Newton-GMRES from the integrated state
    with J v by a finite difference of the time stepper, restart 50,
    tolerance 1e-10 (Newton at dt = 1 as _Newton, Main.py:430-554)."""
    torch, dev = _gpu()
    from spectraldoublediffusiveconvection_amd import newton
    c = C.THIN_GAP
    K, Nr = c2["K"], c2["Nr"]
    X_int = c2["X_end"]
    o1 = _oracle_ops(K, Nr, dt=1.0)
    kr = newton.Krylov(_plan(dev, K, Nr, dt=1.0), restart=50)
    # from the integrated state as C2 runs it, and from that state perturbed
    # by 1% (so that Newton has several iterations to do)
    pert = 1.0 + 0.01 * np.cos(np.arange(X_int.size))
    sols = []
    for name, X_start in (("integrated", X_int), ("perturbed", X_int * pert)):
        start_res = rel(O.step(X_start, o1, c["Ra"], c["Ra_s"]), X_start)
        Xs, info = kr.newton(torch.as_tensor(X_start, device="cuda"), c["Ra"], c["Ra_s"],
                             tol_newton=1e-10, max_newton=10, fd=True)
        Xs = Xs.cpu().numpy()
        res = rel(O.step(Xs, o1, c["Ra"], c["Ra_s"]), Xs)
        print("Newton from the %s state (FD J v, restart 50): %s; residual %.2e -> %.2e"
              % (name, info, start_res, res))
        assert info["converged"], info
        assert info["error"] <= 1e-10
        assert res <= 1e-9
        # the steady state of the dt = 1 stepper is the dt = 5e-3 stepper's too
        assert rel(O.step(Xs, _oracle_ops(K, Nr), c["Ra"], c["Ra_s"]), Xs) <= 1e-11
        sols.append(Xs)
    assert rel(sols[1], sols[0]) <= 1e-9
