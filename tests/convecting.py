"""The convecting regime of the long-run parity tests (helpers, no tests).

The reference's thin-gap validation case (Run_Tests.py:170-196): d = 0.353,
Ra_T = 2360, Ra_S = 0, Pr = 1, Tau = 1.  Its steady convecting state at
N_r = 24, N_fm = 48 is cached in ``golden/convecting_thin_gap.npz``
(written by ``golden/make_convecting.py`` with the oracle) and moved to the
test resolutions by ``cheb_resample``.
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "convecting_thin_gap.npz")
THIN_GAP = dict(d=0.353, Ra=2360.0, Ra_s=0.0, Pr=1.0, Tau=1.0)
FIXTURE_SIZE = (48, 24)          # (N_fm, N_r) of the cached steady state
DT = 5e-3                        # time step of the trajectories
# tag: (N_fm, N_r, steps of the oracle trajectory)
RUNS = {"c3": (512, 40, 2000), "c2": (256, 30, 10000)}


def load():
    return dict(np.load(FIXTURE))


def cheb_resample(X, K0, Nr0, K1, Nr1):
    """A state (reference layout X[f*N + k*nr + i]) moved between resolutions:
    each mode row is the polynomial through its nr0 interior values and the
    zero wall values on the Chebyshev-Gauss-Lobatto points of N_r0 (barycentric
    form), evaluated at the interior points of N_r1; theta modes are zero padded
    or truncated.  Both grids map [-1, 1] to [R_1, R_2] affinely
    (Matrix_Operators.py:10-28), so interpolating in x is interpolating in r."""
    nr0, nr1 = Nr0 - 1, Nr1 - 1
    x0 = np.cos(np.pi * np.arange(Nr0 + 1) / Nr0)
    x1 = np.cos(np.pi * np.arange(1, Nr1) / Nr1)
    w = (-1.0) ** np.arange(Nr0 + 1)
    w[0] *= 0.5
    w[-1] *= 0.5
    diff = x1[:, None] - x0[None, :]
    exact = np.abs(diff) < 1e-15
    diff[exact] = 1.0
    Cm = w[None, :] / diff
    Cm /= Cm.sum(axis=1, keepdims=True)
    for i, j in zip(*np.nonzero(exact)):
        Cm[i, :] = 0.0
        Cm[i, j] = 1.0
    Cm = Cm[:, 1:-1]
    out = np.zeros(3 * K1 * nr1)
    k = min(K0, K1)
    for f in range(3):
        A = np.asarray(X[f * K0 * nr0:(f + 1) * K0 * nr0]).reshape(K0, nr0)
        out[f * K1 * nr1:(f + 1) * K1 * nr1].reshape(K1, nr1)[:k] = A[:k] @ Cm.T
    return out


def projection_vector(n, j):
    """Seeded normal vector j of length n (state checksums)."""
    return np.random.default_rng(100 + j).standard_normal(n)


def start_state(K, Nr, fixture=None):
    """The cached steady state moved to (K, Nr)."""
    z = fixture if fixture is not None else load()
    return cheb_resample(z["X24"], *FIXTURE_SIZE, K, Nr)


def steady_on_device(K, Nr, steps=20000):
    """The cached steady state moved to (K, Nr) and stepped there to steady
    on the GPU (dt = 0.075, symmetric, as the reference runs the case; the
    fixed point of the step does not depend on dt).  Returns a numpy state."""
    import torch

    from spectraldoublediffusiveconvection_amd import device as dev
    c = THIN_GAP
    sp = dev.Plan.build(K, Nr, c["d"], 0.075, c["Pr"], c["Tau"], symmetric=True)
    X = sp.step(torch.as_tensor(start_state(K, Nr), device="cuda"), c["Ra"], c["Ra_s"],
                nsteps=steps)
    out = X[0].cpu().numpy()
    sp.close()
    return out
