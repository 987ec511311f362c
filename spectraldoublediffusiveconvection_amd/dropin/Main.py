"""Drop-in for the reference's Main.py: operator assembly, diagnostics, the
time stepper and the Newton solver, with every step, linearised step and
diagnostic computed by libsddc on the GPU.

Signatures and return values follow Main.py (file:line on each function).
_Newton runs the device Newton-Krylov solver (restarted GMRES in float64 with
CGS2 Gram-Schmidt, sddc_newton) with the reference's tolerances and limits;
set KRYLOV = "lgmres" to use scipy's lgmres on the host around the device's
step and linearised step instead (the reference's own Krylov method).  Both
converge to the same steady state within the Newton tolerance.  The
continuation drivers (_NewtonC, _ContinC, _Continuation) run their Newton and
pseudo-arclength steps on the device (sddc_newton, sddc_arclength_step).
Checkpoints are HDF5 files with the reference's groups, written by h5py when
it is importable and otherwise by the package's own writer (_hdf5.write:
the same groups and datasets, readable by h5py).
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .. import operators as _ops
from . import Matrix_Operators as _mo
from . import _hdf5
from . import _registry as _reg


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64)), device="cuda").view(1, -1)


def _host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()[0]


def Base_State_Coeffs(d):
    """T_0 = -A_T/r + B_T (Main.py:23-38)."""
    return _ops.base_state_coeffs(d)


def _nusselt_ends(T_hat, d, R, D, N_fm, nr):
    N = N_fm * nr
    x = np.zeros(3 * N)
    x[N:2 * N] = np.asarray(T_hat, dtype=np.float64)[:N]
    p = _reg.geometry_plan(D, R, N_fm, False, d=d)
    out = _host(p.diagnostics(_dev(x)))
    return out[2], out[4]


def Nusselt(T_hat, d, R, D, N_fm, nr, check=True):
    """Nu - 1 at R_1; with check, the relative difference to the value at R_2
    must stay below 1e-3 (Main.py:42-68)."""
    nu1, nu2 = _nusselt_ends(T_hat, d, R, D, N_fm, nr)
    error = abs(nu2 - nu1) / abs(nu2)
    if (check is True) and (error > 1e-03):
        raise ValueError("|Nu(R_1) - Nu(R_2)|/|Nu(R_1)| = ", error, " > .1%, Increase Nr \n")
    return nu1


def Kinetic_Energy(X_hat, R, D, N_fm, nr, symmetric):
    """Volume-averaged kinetic energy (Main.py:71-134)."""
    p = _reg.geometry_plan(D, R, N_fm, symmetric)
    return _host(p.diagnostics(_dev(X_hat)))[1]


def Eq_SYM(X1, R):
    """0/1 mask keeping odd psi rows and even T/S rows (Main.py:137-176)."""
    X_sym = np.ones(np.shape(X1))
    nr = len(R[1:-1])
    N_fm = int(np.shape(X1)[0] / (3 * nr))
    N = N_fm * nr
    for k in range(N_fm):
        if k % 2 == 0:
            X_sym[k * nr:k * nr + nr] = 0.0
        else:
            X_sym[N + k * nr:N + k * nr + nr] = 0.0
            X_sym[2 * N + k * nr:2 * N + k * nr + nr] = 0.0
    return X_sym


def Build_Matrix_Operators(N_fm, N_r, d, dt, Pr, Tau):
    """(D, R, Rsq, DT0, gr_K, L_inv_A4, L_inv_NAB2_T, L_inv_NAB2_S, args_A4,
    args_FXS) as Main.py:179-225 returns them (host float64)."""
    ops = _ops.build(N_fm, N_r, d, dt, Pr, Tau)
    _reg.remember_geometry(ops.D, ops.R)
    nr = N_r - 1
    Rsq = _mo.R2(ops.R, N_fm)
    gr_K = _mo.kGR_RT(ops.R, N_fm, d)
    IR2 = np.ascontiguousarray(np.diag(ops.IR2_A4))
    IR4 = np.ascontiguousarray(np.diag(ops.IR4_A4))
    args_A4 = (np.ascontiguousarray(ops.D2_A4), IR4, IR2, N_fm, nr)
    args_FXS = (ops.D, ops.R, N_fm, nr)
    return (ops.D, ops.R, Rsq, np.array(ops.DT0), gr_K, list(ops.Linv_A4), list(ops.Linv_T),
            list(ops.Linv_S), args_A4, args_FXS)


def _write_checkpoint(save_filename, X_DATA, scalars, params):
    try:
        import h5py
    except ImportError:
        h5py = None
    if h5py is not None:
        with h5py.File(save_filename, "w") as f:
            f.create_group("Checkpoints")["X_DATA"] = X_DATA
            g = f.create_group("Scalar_Data")
            for k, v in scalars.items():
                g[k] = v
            g = f.create_group("Parameters")
            for k, v in params.items():
                g[k] = v
        return save_filename
    return _hdf5.write(save_filename, {"Checkpoints": {"X_DATA": X_DATA},
                                       "Scalar_Data": dict(scalars),
                                       "Parameters": dict(params)})


def _Time_Step(X, Ra, Ra_s, Tau, Pr, d, N_fm, N_r, symmetric, save_filename='TimeStep_0.h5',
               start_time=0, Total_time=1, dt=1e-04, linear=False, Verbose=False):
    """Integrate Total_time/dt steps of Step_Python with the loop's symmetry
    mask, diagnostics every step and a checkpoint every tenth of the run
    (Main.py:228-329).  Returns the last state."""
    plan = _reg.step_plan(N_fm, N_r, d, dt, Pr, Tau, symmetric)
    N_iters = int(Total_time / dt)
    N_save = max(int(N_iters / 10), 1)
    N_print = 10
    x = _dev(X)
    xn = torch.empty_like(x)
    Time, X_DATA, KE, NuT, NuS, Norm = [], [], [], [], [], []
    params = {"Ra": Ra, "Ra_s": Ra_s, "Tau": Tau, "Pr": Pr, "d": d, "N_r": N_r, "N_fm": N_fm,
              "dt": dt, "start_time": start_time, "symmetric": symmetric}
    timer = 0.0
    X_new = np.asarray(X, dtype=np.float64)
    for iteration in range(N_iters):
        torch.cuda.synchronize()
        st = time.time()
        plan.step(x, Ra, Ra_s, nsteps=1, linear=linear, out=xn)
        torch.cuda.synchronize()
        end = time.time()
        dg = _host(plan.diagnostics(xn))
        Norm.append(dg[0])
        KE.append(dg[1])
        NuT.append(dg[2])
        NuS.append(dg[3])
        Time.append(start_time + dt * iteration)
        if (iteration % N_print == 0) and (Verbose is True):
            print("Iteration =%i, Time = %e, Energy = %e, NuT = %e \n " %
                  (iteration, start_time + dt * iteration, KE[-1], NuT[-1]))
        if iteration % N_save == 0:
            X_DATA.append(_host(xn))
            _write_checkpoint(save_filename, X_DATA, {"Norm": Norm, "KE": KE, "Nu_T": NuT,
                                                      "Nu_S": NuS, "Time": Time}, params)
        x, xn = xn, x   # the step already applied the symmetry mask
        if iteration > 0:
            timer += end - st
    if N_iters > 1:
        print("Avg time = %e \n" % (timer / (N_iters - 1)))
    if N_iters > 0:
        X_new = _host(x)
    return X_new


KRYLOV = "device"
_krylov = {}


def _krylov_for(plan, restart):
    """One Krylov basis, rebuilt when the plan (an object the registry may
    have closed and replaced) or the restart length changes."""
    from ..newton import Krylov
    kr = _krylov.get("k")
    if kr is None or kr.plan is not plan or kr.restart != int(restart):
        if kr is not None:
            kr.close()
        kr = _krylov["k"] = Krylov(plan, restart=restart)
    return kr


def _Newton(X, Ra, Ra_s, Tau, Pr, d, N_fm, N_r, symmetric, dt=1, tol_newton=1e-08,
            tol_gmres=1e-04, Krylov_Space_Size=300):
    """Steady state by Newton iteration on Step(X) - X (Main.py:430-554).
    Returns (X, Norm, KE, NuT, NuS, True) or (None, None, None, None, None, False)."""
    if KRYLOV == "lgmres":
        return _newton_lgmres(X, Ra, Ra_s, Tau, Pr, d, N_fm, N_r, symmetric, dt, tol_newton,
                              tol_gmres, Krylov_Space_Size)
    plan = _reg.step_plan(N_fm, N_r, d, dt, Pr, Tau, symmetric)
    kr = _krylov_for(plan, Krylov_Space_Size)
    Xd, info = kr.newton(_dev(X), Ra, Ra_s, tol_newton=tol_newton, tol_gmres=tol_gmres,
                         max_newton=5, max_restarts=250)
    print('Newton Iterations = %d, Error = %e, Krylov iterations = %d'
          % (info["iterations"], info["error"], info["krylov_iterations"]), "\n")
    if not info["converged"]:
        print('Newton iteration could not converge')
        return None, None, None, None, None, False
    ops = plan.ops
    R, D = ops.R, ops.D
    nr = N_r - 1
    N = N_fm * nr
    X = _host(Xd.view(1, -1))
    Norm = np.linalg.norm(X, 2)
    KE = Kinetic_Energy(X, R, D, N_fm, nr, symmetric)
    NuT = Nusselt(X[N:2 * N], d, R, D, N_fm, nr)
    NuS = Nusselt(X[2 * N:3 * N], d, R, D, N_fm, nr)
    return X, Norm, KE, NuT, NuS, True


def _newton_lgmres(X, Ra, Ra_s, Tau, Pr, d, N_fm, N_r, symmetric, dt, tol_newton, tol_gmres,
                   Krylov_Space_Size):
    """_Newton with scipy's lgmres on the host (the reference's Krylov method)."""
    import scipy.sparse.linalg as spla
    plan = _reg.step_plan(N_fm, N_r, d, dt, Pr, Tau, symmetric)
    ops = plan.ops
    R, D = ops.R, ops.D
    nr = N_r - 1
    N = N_fm * nr
    X = np.asarray(X, dtype=np.float64)
    X_SYM = Eq_SYM(X, R) if symmetric is True else 1
    error, iteration, exit = 1.0, 0, 1.0
    while (error > tol_newton) and (iteration < 5):
        X = X_SYM * X
        xd = _dev(X)
        fx = _host(plan.step(xd, Ra, Ra_s)) - X

        def DF(dv, xd=xd):
            dv = np.asarray(dv, dtype=np.float64).reshape(-1)
            return _host(plan.step_linearized(_dev(dv), xd, Ra, Ra_s)) - dv

        dfx = spla.LinearOperator((X.shape[0], X.shape[0]), matvec=DF, dtype='float64')
        b_norm = np.linalg.norm(fx, 2)
        dv, exit = spla.lgmres(dfx, fx, maxiter=250, inner_m=Krylov_Space_Size, atol=tol_gmres * b_norm)
        X = X - dv
        error = np.linalg.norm(dv, 2) / np.linalg.norm(X, 2)
        print('Newton Iteration = %d, Error = %e' % (iteration, error), "\n")
        iteration += 1
    if (iteration == 5) or (exit != 0):
        print('Newton iteration could not converge')
        return None, None, None, None, None, False
    Norm = np.linalg.norm(X, 2)
    KE = Kinetic_Energy(X, R, D, N_fm, nr, symmetric)
    NuT = Nusselt(X[N:2 * N], d, R, D, N_fm, nr)
    NuS = Nusselt(X[2 * N:3 * N], d, R, D, N_fm, nr)
    return X, Norm, KE, NuT, NuS, True


# --------------------------------------------------------------------------
# continuation in Ra_T (Main.py:677-1045)
# --------------------------------------------------------------------------

class result():
    """Record of a continuation run (Main.py:677-713)."""

    def __init__(self):
        self.NuT = []
        self.NuS = []
        self.KE = []
        self.Norm = []
        self.Ra = []
        self.Ra_dot = []
        self.Y_FOLD = []
        self.X_DATA = []
        self.Ra_DATA = []
        self.Iterations = 0

    def __str__(self):
        i = self.Iterations
        return ('Continuation succeed \n'
                + 'Total iterations  = ' + str(i) + '\n'
                + 'Ra                = ' + str(self.Ra[i]) + '\n'
                + 'Ra_dot            = ' + str(self.Ra_dot[i]) + '\n'
                + 'NuT               = ' + str(self.NuT[i]) + '\n'
                + 'NuS               = ' + str(self.NuS[i]) + '\n'
                + 'KE                = ' + str(self.KE[i]) + '\n')


def _NewtonC(Y, sign, ds, **kwargs_f):
    """Natural-parameter step: Newton at Ra_T + sign*ds from Y = (X, Ra_T);
    ds doubles on success and halves on failure (Main.py:717-739)."""
    X = Y[0:-1].copy()
    Ra = Y[-1].copy() + sign * ds
    kwargs_f["Ra"] = Ra
    X_new, Norm, KE, NuT, NuS, exitcode = _Newton(X, **kwargs_f, tol_newton=1e-08)
    if exitcode is True:
        ds *= 2
        return np.hstack((X_new, Ra)), sign, ds, Norm, KE, NuT, NuS, exitcode
    ds *= 0.5
    return Y, sign, ds, Norm, KE, NuT, NuS, exitcode


FINITE_DIFFERENCE = False   # J v of the continuation by finite differences of the step


def _ContinC(Y_0_dot, Y_0, sign, ds, Ra, Ra_s, Tau, Pr, d, N_fm, N_r, symmetric, dt=1,
             tol_newton=1e-08, tol_gmres=1e-04, Krylov_Space_Size=300):
    """One pseudo-arclength step from Y_0 = (X, Ra_T) (Main.py:742-950):
    returns (Y_dot, Y, sign, ds, Norm, KE, NuT, NuS, True) or raises
    ValueError as the reference does.  The predictor, the corrector on the
    bordered system and the tangent run on the device (sddc_arclength_step)."""
    from .._lib import ARC_CORRECTOR_FAILED, ARC_PREDICTOR_FAILED, ARC_STEP_TOO_SMALL
    plan = _reg.step_plan(N_fm, N_r, d, dt, Pr, Tau, symmetric)
    kr = _krylov_for(plan, Krylov_Space_Size)
    Yd, Ydd, ds, info = kr.arclength_step(_dev(Y_0), _dev(Y_0_dot), Ra_s, sign, ds,
                                          tol_newton=tol_newton, tol_gmres=tol_gmres,
                                          fd=FINITE_DIFFERENCE)
    for _ in range(info["halvings"]):
        print('Reducing the step-size ds by half \n')
    if info["status"] == ARC_PREDICTOR_FAILED:
        raise ValueError("\n LGMRES couldn't converge when predicting ξ \n")
    if info["status"] == ARC_STEP_TOO_SMALL:
        raise ValueError("\n Step-size ds has become too small terminating \n")
    if info["status"] == ARC_CORRECTOR_FAILED:
        raise ValueError("\n LGMRES couldn't converge when predicting dY \n")
    print('Psuedo-Arc Iterations = %d, Error X = %e, Error µ = %e'
          % (info["iterations"], info["error_X"], info["error_mu"]), "\n")
    ops = plan.ops
    R, D = ops.R, ops.D
    nr = N_r - 1
    N = N_fm * nr
    Y = _host(Yd.view(1, -1))
    Y_dot = _host(Ydd.view(1, -1))
    X = Y[0:-1].copy()
    Norm = np.linalg.norm(X, 2)
    KE = Kinetic_Energy(X, R, D, N_fm, nr, symmetric)
    NuT = Nusselt(X[N:2 * N], d, R, D, N_fm, nr)
    NuS = Nusselt(X[2 * N:3 * N], d, R, D, N_fm, nr)
    return Y_dot, Y, sign, ds, Norm, KE, NuT, NuS, True


def _write_continuation(filename, Result, params):
    """Checkpoints, Parameters and Bifurcation groups (Main.py:1024-1040)."""
    try:
        import h5py
    except ImportError:
        h5py = None
    if h5py is not None:
        with h5py.File(filename, "w") as f:
            g = f.create_group("Checkpoints")
            g["X_DATA"] = Result.X_DATA
            g["Ra_DATA"] = Result.Ra_DATA
            g = f.create_group("Parameters")
            for k, v in params.items():
                g[k] = v
            g = f.create_group("Bifurcation")
            for k, v in vars(Result).items():
                g.create_dataset(k, data=v)
        return filename
    return _hdf5.write(filename, {"Checkpoints": {"X_DATA": Result.X_DATA,
                                                  "Ra_DATA": Result.Ra_DATA},
                                  "Parameters": dict(params),
                                  "Bifurcation": dict(vars(Result))})


def _Continuation(filename, N_steps, sign, Y, **kwargs):
    """Branch continuation in Ra_T from Y = (X, Ra_T) for N_steps points
    (Main.py:958-1045): natural-parameter Newton steps while ds > 1, switching
    to pseudo-arclength below, fold detection by a sign change of Ra_dot, and a
    checkpoint every fifth point.  Returns the result record (the reference
    returns None after writing the same record)."""
    Y = Y.copy()
    Y_dot = 0. * Y
    ds, ds_min, ds_max = 0.01, 1.0, 10.0
    Result = result()
    Y_dot_new = None   # the last pseudo-arclength tangent, kept across Newton steps
    while Result.Iterations < N_steps:
        if ds > ds_min:
            OUT = _NewtonC(Y, sign, ds, **kwargs)
            if OUT[-1] is False:
                print('Switching to arc-length \n')
                Y_dot_new, Y_new, sign, ds, Norm, KE, NuT, NuS, exitcode = _ContinC(
                    Y_dot, Y, sign, ds, **kwargs)
            else:
                Y_new, sign, ds, Norm, KE, NuT, NuS = OUT[:7]
            if ds > ds_max:
                ds = ds_max
        else:
            Y_dot_new, Y_new, sign, ds, Norm, KE, NuT, NuS, exitcode = _ContinC(
                Y_dot, Y, sign, ds, **kwargs)
            if Y_dot_new[-1] * Y_dot[-1] < 0:   # fold (Main.py:995-997)
                Result.Y_FOLD.append(Y_new)
            sign = 1.0 if Y_new[-1] > Y[-1] else -1.0
        Y = Y_new.copy()
        Y_dot = Y_dot_new.copy() if Y_dot_new is not None else 0. * Y
        Result.Ra.append(Y[-1])
        Result.Ra_dot.append(Y_dot[-1])
        Result.Norm.append(Norm)
        Result.KE.append(KE)
        Result.NuT.append(NuT)
        Result.NuS.append(NuS)
        if Result.Iterations % 5 == 0:
            Result.X_DATA.append(Y[0:-1])
            Result.Ra_DATA.append(Y[-1])
            _write_continuation(filename, Result, kwargs)
        print(Result, flush=True)
        Result.Iterations += 1
    return Result
