"""Write the convecting-regime fixture of the long-run parity tests.

    python tests/golden/make_convecting.py

The reference's thin-gap validation case (Run_Tests.py:170-196): d = 0.353,
Ra_T = 2360, Ra_S = 0, Pr = 1, Tau = 1, N_r = 24, N_fm = 48, dt = 0.075,
equatorially symmetric, T = 2000 (26667 steps) from a seeded small state.
The oracle's run ends on a steady convecting state (||X|| ~ 1.5,
||dt F(X)|| / ||X|| ~ 3.4 at dt = 0.075).

From that state, moved to N_r = 40, N_fm = 512 (C3) and N_r = 30, N_fm = 256
(C2) by ``convecting.cheb_resample``, the oracle runs plain float64 steps of
dt = 5e-3 with symmetric = False: 2000 steps at C3 and 10^4 steps at C2.
Every 100 steps it records ||X||, the kinetic energy and Nu - 1 of T and S
at R_1 (Main.py:292-295); every 500 steps ||dt F(X)|| / ||X||; at the end
the state's dot products with four seeded normal vectors.  The GPU tests
compare their trajectories from the same start against these numbers.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, TESTS)

from oracle import sddc_oracle as O  # noqa: E402
import convecting as C  # noqa: E402


def steady_thin_gap():
    c = C.THIN_GAP
    K, Nr = C.FIXTURE_SIZE
    n3 = 3 * K * (Nr - 1)
    X = np.random.default_rng(8).random(n3)
    X = 1e-3 * X / np.linalg.norm(X)
    o = O.build_operators(K, Nr, c["d"], 0.075, c["Pr"], c["Tau"])
    for _ in range(26667):
        X = O.step(X, o, c["Ra"], c["Ra_s"], symmetric=True)
    res = np.linalg.norm(O.step(X, o, c["Ra"], c["Ra_s"], symmetric=True) - X) / np.linalg.norm(X)
    assert res < 1e-13, res
    return X


def trajectory(X24, K, Nr, nsteps):
    c = C.THIN_GAP
    nr = Nr - 1
    o = O.build_operators(K, Nr, c["d"], C.DT, c["Pr"], c["Tau"])
    X = C.cheb_resample(X24, *C.FIXTURE_SIZE, K, Nr)
    diag, ratio = [], []
    t0 = time.time()
    for n in range(nsteps + 1):
        if n % 100 == 0:
            N = K * nr
            diag.append([np.linalg.norm(X), O.kinetic_energy(X, o.R, o.D, K, nr),
                         O.nusselt(X[N:2 * N], c["d"], o.R, o.D, K, nr, check=False),
                         O.nusselt(X[2 * N:], c["d"], o.R, o.D, K, nr, check=False)])
        if n % 500 == 0:
            F = O.nlin_fx(X, o)
            ratio.append(C.DT * np.linalg.norm(F) / np.linalg.norm(X))
            print("  %d/%d steps, |X| %.6f, |dt F|/|X| %.4f, %.0f s" % (
                n, nsteps, np.linalg.norm(X), ratio[-1], time.time() - t0), flush=True)
        if n < nsteps:
            X = O.step(X, o, c["Ra"], c["Ra_s"])
    proj = np.array([C.projection_vector(3 * K * nr, j) @ X for j in range(4)])
    return np.array(diag), np.array(ratio), proj


def main():
    X24 = steady_thin_gap()
    out = {"X24": X24}
    for tag, (K, Nr, nsteps) in C.RUNS.items():
        print(tag, K, Nr, nsteps, flush=True)
        diag, ratio, proj = trajectory(X24, K, Nr, nsteps)
        out[tag + "_diag"] = diag
        out[tag + "_ratio"] = ratio
        out[tag + "_proj"] = proj
    np.savez(C.FIXTURE, **out)
    print("wrote", C.FIXTURE)


if __name__ == "__main__":
    main()
