"""Write golden input/output vectors from the reference implementation.

Run in the development container only (the reference is not present on GPU
machines):

    python tests/golden/make_golden.py /root/reference

The reference's JIT decorator (numba) and HDF5 writer (h5py) are absent here,
and NumPy 2 removed ``np.RankWarning``; a shim installed in ``sys.modules``
before the import supplies an identity ``njit``, ``numba.typed.List = list``,
``np.RankWarning`` and an in-memory h5py that records what ``_Time_Step``
writes.  The JIT bodies then run as plain NumPy.

Vectors are written only for the functions no closed form pins:
NLIN_FX, NLIN_DFX, Derivatives, the composition of the step inside
``_Time_Step`` (Main.py:255-283, run through the real driver) and
Kinetic_Energy (Main.py:71-134).
"""
from __future__ import annotations

import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def install_shims():
    numba = types.ModuleType("numba")

    def njit(*args, **kwargs):
        if len(args) == 1 and callable(args[0]) and not kwargs:
            return args[0]
        return lambda f: f

    numba.njit = njit
    typed = types.ModuleType("numba.typed")
    typed.List = list
    numba.typed = typed
    sys.modules["numba"] = numba
    sys.modules["numba.typed"] = typed
    if not hasattr(np, "RankWarning"):
        np.RankWarning = np.exceptions.RankWarning
    if not hasattr(np, "trapz"):
        np.trapz = np.trapezoid

    h5 = types.ModuleType("h5py")
    h5.files = {}

    class Group(dict):
        def create_group(self, name):
            g = Group()
            self[name] = g
            return g

        def create_dataset(self, name, data=None):
            self[name] = data

    class File(Group):
        def __init__(self, name, mode="r", **kw):
            super().__init__()
            h5.files[name] = self

        def close(self):
            pass

    h5.File = File
    sys.modules["h5py"] = h5
    import matplotlib
    matplotlib.use("Agg")
    return h5


def main(ref_root):
    h5 = install_shims()
    sys.path.insert(0, ref_root)
    import Matrix_Operators as MO  # noqa: E402
    import Main  # noqa: E402

    out = {}
    # Workload: N_r=20 (nr=19), N_fm=64, d=0.35, Pr=1, Tau=1/15, Ra_s=500,
    # Ra=9000, dt=5e-3 -- the CPU-runnable configuration.
    N_r, N_fm, d, Pr, Tau, Ra, Ra_s, dt = 20, 64, 0.35, 1.0, 1.0 / 15.0, 9000.0, 500.0, 5e-3
    nr = N_r - 1
    N = nr * N_fm
    D, R = MO.cheb_radial(N_r, d)
    rng = np.random.default_rng(20240611)
    X_small = 1e-3 * rng.standard_normal(3 * N)
    X_big = 0.5 * rng.standard_normal(3 * N)
    dv = 0.3 * rng.standard_normal(3 * N)
    out.update(N_r=N_r, N_fm=N_fm, d=d, Pr=Pr, Tau=Tau, Ra=Ra, Ra_s=Ra_s, dt=dt,
               X_small=X_small, X_big=X_big, dv=dv)

    for sym in (False, True):
        tag = "sym" if sym else "full"
        out["nlin_fx_big_" + tag] = MO.NLIN_FX(X_big, D, R, N_fm, nr, sym)
        out["nlin_fx_small_" + tag] = MO.NLIN_FX(X_small, D, R, N_fm, nr, sym)
        out["nlin_dfx_" + tag] = MO.NLIN_DFX(dv, X_big, D, R, N_fm, nr, sym)
        out["ke_big_" + tag] = Main.Kinetic_Energy(X_big, R, D, N_fm, nr, sym)

    # Derivatives' nine slabs for X_big (Matrix_Operators.py:630-740)
    JPSI = MO.J_theta_RT(X_big[0:N], nr, N_fm, False)
    OMEGA = MO.A2_SINE_R2(X_big[0:N], N_fm, nr, D, R, False)
    slabs = MO.Derivatives(X_big, np.ascontiguousarray(JPSI), np.ascontiguousarray(OMEGA),
                           np.ascontiguousarray(D[1:-1, 1:-1]), N_fm, nr, False)
    names = ["JT", "kDpsi", "komega", "DT", "DC", "omega", "Dpsi", "kT", "kC"]
    for name, s in zip(names, slabs):
        out["slab_" + name] = s

    # Full driver: n steps through _Time_Step (Main.py:228-329); the in-memory
    # h5py captures the per-step diagnostics it writes.
    def run_driver(X0, n, sym, linear, Ra_, Ra_s_, dt_, N_r_, N_fm_, d_, Pr_, Tau_, fname):
        total = n * dt_
        assert int(total / dt_) == n, (total, dt_, n)
        Xn = Main._Time_Step(X0.copy(), Ra_, Ra_s_, Tau_, Pr_, d_, N_fm_, N_r_, sym,
                             save_filename=fname, start_time=0, Total_time=total, dt=dt_,
                             linear=linear, Verbose=False)
        f = h5.files[fname]
        return Xn, f["Scalar_Data"]

    for sym in (False, True):
        tag = "sym" if sym else "full"
        Xn, sd = run_driver(X_small, 20, sym, False, Ra, Ra_s, dt, N_r, N_fm, d, Pr, Tau,
                            "drv_%s.h5" % tag)
        out["traj20_X_" + tag] = Xn
        for key in ("Norm", "KE", "Nu_T", "Nu_S"):
            out["traj20_%s_%s" % (key, tag)] = np.asarray(sd[key])
    # One step from a large-amplitude state (nonlinear term dominant)
    Xn, sd = run_driver(X_big, 10, False, False, Ra, Ra_s, dt, N_r, N_fm, d, Pr, Tau, "drv_big.h5")
    out["traj10big_X_full"] = Xn
    out["traj10big_Norm_full"] = np.asarray(sd["Norm"])
    out["traj10big_KE_full"] = np.asarray(sd["KE"])
    # linear run (nonlinear term off, Main.py:264)
    Xn, sd = run_driver(X_small, 10, False, True, Ra, Ra_s, dt, N_r, N_fm, d, Pr, Tau, "drv_lin.h5")
    out["traj10lin_X_full"] = Xn

    path = os.path.join(HERE, "golden_c1.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "keys:", len(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
