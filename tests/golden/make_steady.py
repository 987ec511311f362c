"""Write the steady convective state the continuation tests start from.

    python tests/golden/make_steady.py

The reference's wide-gap steady-convection case (Run_Tests.py:140-235):
d = 2, N_fm = 32, N_r = 16, Ra_T = 6780, Ra_S = 0, Pr = 10, Tau = 1,
equatorially symmetric.  The oracle time-steps a seeded small state for
26667 steps of dt = 0.075, then its Newton solve (lgmres, as _Newton) at
dt = 1 converges the steady state to 1e-10.  The fixture caches that state
so that tests do not repeat the half-minute time integration.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import sddc_oracle as O  # noqa: E402

CASE = dict(d=2.0, N_fm=32, N_r=16, Ra=6780.0, Ra_s=0.0, Tau=1.0, Pr=10.0)


def main():
    c = CASE
    n3 = 3 * c["N_fm"] * (c["N_r"] - 1)
    X = np.random.default_rng(6).random(n3)
    X = 1e-3 * X / np.linalg.norm(X)
    o = O.build_operators(c["N_fm"], c["N_r"], c["d"], 0.075, c["Pr"], c["Tau"])
    for _ in range(26667):
        X = O.step(X, o, c["Ra"], c["Ra_s"], symmetric=True)
    Xs, Norm, KE, NuT, NuS, ok = O.newton(X, c["Ra"], c["Ra_s"], c["Tau"], c["Pr"], c["d"],
                                          c["N_fm"], c["N_r"], True, tol_newton=1e-10)
    assert ok, "Newton did not converge"
    np.savez(os.path.join(HERE, "steady_widegap.npz"), X=Xs, KE=KE, NuT=NuT, NuS=NuS,
             **{k: np.float64(v) for k, v in c.items()})
    print("KE = %.12e  NuT = %.12e  ||X|| = %.12e" % (KE, NuT, Norm))


if __name__ == "__main__":
    main()
