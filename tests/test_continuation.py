"""Pseudo-arclength continuation in Ra_T (_ContinC, Main.py:742-950).

CPU: the oracle's dF/dRa against a difference quotient of the step (which is
affine in Ra_T), and the oracle's continuation step at solution level
(F(X, Ra) = 0, the tangent solves DG Y_dot = e_last) from the cached steady
state of the reference's wide-gap case (tests/golden/make_steady.py).

GPU: the device step (sddc_arclength_step) against the oracle's from the
same state with the same tolerances, and the drop-in _Continuation driver
for three branch points."""
import os

import numpy as np
import pytest

from oracle import sddc_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def case():
    z = np.load(os.path.join(HERE, "golden", "steady_widegap.npz"))
    c = {k: float(z[k]) for k in ("d", "Pr", "Tau", "Ra", "Ra_s")}
    c["N_fm"], c["N_r"] = int(z["N_fm"]), int(z["N_r"])
    c["X"] = z["X"]
    c["o1"] = O.build_operators(c["N_fm"], c["N_r"], c["d"], 1.0, c["Pr"], c["Tau"])
    return c


@pytest.mark.parametrize("sym", [False, True])
def test_pdfmu_is_the_ra_derivative_of_the_step(sym):
    K, Nr = 8, 9
    o = O.build_operators(K, Nr, 0.35, 5e-3, 1.0, 1 / 15.0)
    X = 1e-2 * np.random.default_rng(3).standard_normal(3 * K * (Nr - 1))
    h = 1000.0
    dq = (O.step(X, o, 3000.0 + h, 500.0, symmetric=sym)
          - O.step(X, o, 3000.0, 500.0, symmetric=sym)) / h
    d = O.pdfmu(X, o, sym)
    assert rel(d, dq) < 1e-8
    N = K * (Nr - 1)
    assert not d[N:].any()


def test_steady_state_fixture_is_steady(case):
    X = case["X"]
    F = O.step(X, case["o1"], case["Ra"], case["Ra_s"], symmetric=True) - X
    assert np.linalg.norm(F) / np.linalg.norm(X) < 1e-10
    assert np.linalg.norm(X) > 0.1   # convecting, not the conduction state


def test_oracle_contin_step_solution(case):
    o1 = case["o1"]
    Y0 = np.concatenate((case["X"], [case["Ra"]]))
    Yd, Y, sign, ds = O.contin_step(0 * Y0, Y0, 1.0, 10.0, o1, case["Ra_s"], symmetric=True,
                                    rtol=1e-12)
    X1, mu1 = Y[:-1], Y[-1]
    F = O.step(X1, o1, mu1, case["Ra_s"], symmetric=True) - X1
    assert np.linalg.norm(F) / np.linalg.norm(X1) < 1e-9
    assert 6785.0 < mu1 < 6795.0 and ds == 20.0   # converged within 4 iterations: ds doubled
    delta = 1.0 / (len(Y) - 1)
    assert abs(delta * Yd[:-1] @ Yd[:-1] + (1 - delta) * Yd[-1] ** 2 - 1.0) < 1e-12
    top = (O.step_linearized(Yd[:-1], X1, o1, mu1, case["Ra_s"], True) - Yd[:-1]
           + O.pdfmu(X1, o1, True) * Yd[-1])
    assert np.linalg.norm(top) / np.linalg.norm(Yd[:-1]) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("fd", [False, True])
def test_device_arclength_step_matches_oracle(case, fd):
    import torch

    from spectraldoublediffusiveconvection_amd import _lib, device as dev, newton
    plan = dev.Plan.build(case["N_fm"], case["N_r"], case["d"], 1.0, case["Pr"], case["Tau"],
                          symmetric=True)
    kr = newton.Krylov(plan, restart=300)
    Y0 = np.concatenate((case["X"], [case["Ra"]]))
    T = lambda a: torch.as_tensor(a, device="cuda")
    Y, Yd, ds, info = kr.arclength_step(T(Y0), T(0 * Y0), case["Ra_s"], 1.0, 10.0, rtol=1e-12,
                                        fd=fd)
    assert info["status"] == _lib.ARC_OK and info["tangent_ok"], info
    Y, Yd = Y.cpu().numpy(), Yd.cpu().numpy()
    Ydr, Yr, _, dsr = O.contin_step(0 * Y0, Y0, 1.0, 10.0, case["o1"], case["Ra_s"],
                                    symmetric=True, rtol=1e-12)
    assert ds == dsr
    assert abs(Y[-1] - Yr[-1]) / Yr[-1] < 1e-10
    assert rel(Y[:-1], Yr[:-1]) < 1e-8
    assert abs(Yd[-1] - Ydr[-1]) < 1e-6 and rel(Yd[:-1], Ydr[:-1]) < 1e-3
    F = O.step(Y[:-1], case["o1"], Y[-1], case["Ra_s"], symmetric=True) - Y[:-1]
    assert np.linalg.norm(F) / np.linalg.norm(Y[:-1]) < 1e-9


@pytest.mark.gpu
def test_device_arclength_failure_is_reported(case):
    """A predictor solve that does not converge (forced from the host: a
    two-vector basis, one cycle, a tight tolerance) is the reference's
    'LGMRES couldn't converge when predicting xi' case; the device step
    reports it as a status and leaves Y untouched beyond the prediction."""
    import torch

    from spectraldoublediffusiveconvection_amd import _lib, device as dev, newton
    plan = dev.Plan.build(case["N_fm"], case["N_r"], case["d"], 1.0, case["Pr"], case["Tau"],
                          symmetric=True)
    kr = newton.Krylov(plan, restart=2)
    Y0 = np.concatenate((case["X"], [case["Ra"]]))
    T = lambda a: torch.as_tensor(a, device="cuda")
    _, _, ds, info = kr.arclength_step(T(Y0), T(0 * Y0), case["Ra_s"], 1.0, 10.0, rtol=1e-12,
                                       max_restarts=1)
    assert info["status"] == _lib.ARC_PREDICTOR_FAILED, info
    assert ds == 10.0


@pytest.mark.gpu
def test_dropin_continuation_driver(case, tmp_path):
    from spectraldoublediffusiveconvection_amd.dropin import Main
    Y0 = np.concatenate((case["X"], [case["Ra"]]))
    kw = {"Ra": case["Ra"], "Ra_s": case["Ra_s"], "Tau": case["Tau"], "Pr": case["Pr"],
          "d": case["d"], "N_fm": case["N_fm"], "N_r": case["N_r"], "symmetric": True}
    fn = str(tmp_path / "cont.h5")
    res = Main._Continuation(fn, 3, 1.0, Y0, **kw)
    assert res.Iterations == 3 and len(res.Ra) == 3
    assert case["Ra"] < res.Ra[0] < res.Ra[1] < res.Ra[2]
    assert os.path.exists(fn)
    # every branch point is a steady state
    out = Main._ContinC(0 * Y0, Y0, 1.0, 0.01, **kw)
    assert len(out) == 9 and out[-1] is True
    Y1 = out[1]
    F = O.step(Y1[:-1], case["o1"], Y1[-1], case["Ra_s"], symmetric=True) - Y1[:-1]
    assert np.linalg.norm(F) / np.linalg.norm(Y1[:-1]) < 1e-9
    assert abs(Y1[-1] - res.Ra[0]) / res.Ra[0] < 1e-10
