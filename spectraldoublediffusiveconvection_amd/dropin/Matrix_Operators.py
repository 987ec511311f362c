"""Drop-in for the reference's Matrix_Operators.py.

Per-step operators (the theta recurrences, Derivatives, NLIN_FX, NLIN_DFX and
the two back substitutions) run in libsddc on the GPU.  Operator builders
(cheb_radial, Nabla2/4, R2, kGR_RT, the per-mode inverses), the layout
helpers and INTERP_RADIAL run on the host in float64 NumPy, as in the
reference: they are setup, not steps.  Signatures follow the reference
(file:line on each function); every function returns new arrays.

Notes on the boundary:
  * Derivatives(X_hat, JPSI, OMEGA, Dr, ...) recomputes JPSI and OMEGA from
    X_hat on the device; every reference caller passes J_theta_RT(psi) and
    A2_SINE_R2(psi) of the same X_hat (Matrix_Operators.py:759-760, 827-830).
    Dr must come from a geometry built by cheb_radial in this process.
  * The superseded NAB2_BSub_TSTEP / A4_BSub_TSTEP (Matrix_Operators.py:249,
    328) are not provided; the drivers call the _V2 versions.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import operators as _ops
from . import _registry as _reg


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64)), device="cuda")


def _host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _psi_state(g, N):
    x = np.zeros(3 * N)
    x[:N] = np.asarray(g, dtype=np.float64)[:N]
    return x


# ------------------------------------------------------------- operator setup
def cheb_radial(N, d):
    """Chebyshev-Gauss-Lobatto points on [R_1, R_2] and the differentiation
    matrix (Matrix_Operators.py:10-28)."""
    D, R = _ops.cheb_radial(N, d)
    _reg.remember_geometry(D, R)
    return D, R


def Nabla2(D, r):
    """Matrix_Operators.py:32-45."""
    return _ops.nabla2(D, r)


def Nabla4(D, r):
    """Matrix_Operators.py:49-76."""
    return _ops.nabla4(D, r)


def R2(R, N_fm):
    """Block-diagonal r^2 over the cosine modes, CSR (Matrix_Operators.py:79-94)."""
    from scipy.sparse import block_diag, diags
    GR = diags(np.asarray(R)[1:-1] ** 2, 0, format="csr")
    return block_diag([GR] * N_fm, format="csr")


def kGR_RT(R, N_fm, d):
    """Buoyancy coupling of psi row jj to T/S row jj+1 with -(jj+1) R_1^2/r^2,
    CSR (Matrix_Operators.py:97-128)."""
    from scipy.sparse import bmat, diags
    R = np.asarray(R)
    nr = len(R[1:-1])
    R_1 = 1.0 / d
    GR = diags((R_1 ** 2) / (R[1:-1] ** 2), 0, format="csr")
    Z = diags(0.0 * np.ones(nr), 0, format="csr")
    blocks = [[(-k * GR if k == jj + 1 else Z) for k in range(N_fm)] for jj in range(N_fm)]
    return bmat(blocks, format="csr")


def NAB2_TSTEP_MATS(dt, N_fm, nr, D, R):
    """Per-mode inverses of r^2(1 - dt Lap), index jj (Matrix_Operators.py:1014-1030)."""
    return list(_ops.nab2_inverses(dt, N_fm, nr, D, R))


def A4_TSTEP_MATS(dt, N_fm, nr, D, R):
    """Per-mode inverses of the psi operator (Matrix_Operators.py:1089-1112)."""
    return list(_ops.a4_inverses(dt, N_fm, nr, D, R))


# -------------------------------------------------------------- layout helpers
def Vecs_to_X(PSI, T, C, N_fm, nr, symmetric):
    """(nr, N_fm) x 3 -> mode-major state (Matrix_Operators.py:530-576)."""
    N = N_fm * nr
    NX = np.zeros(3 * N)
    psi_rows = range(1, N_fm, 2) if symmetric else range(N_fm)
    ts_rows = range(0, N_fm, 2) if symmetric else range(N_fm)
    for ii in psi_rows:
        NX[ii * nr:(ii + 1) * nr] = PSI[:, ii]
    for ii in ts_rows:
        NX[N + ii * nr:N + (ii + 1) * nr] = T[:, ii]
        NX[2 * N + ii * nr:2 * N + (ii + 1) * nr] = C[:, ii]
    return NX


def X_to_Vecs(X, N_fm, nr, symmetric):
    """Mode-major state -> (nr, N_fm) x 3 (Matrix_Operators.py:579-627)."""
    N = N_fm * nr
    PSI, T, C = np.zeros((nr, N_fm)), np.zeros((nr, N_fm)), np.zeros((nr, N_fm))
    psi_rows = range(1, N_fm, 2) if symmetric else range(N_fm)
    ts_rows = range(0, N_fm, 2) if symmetric else range(N_fm)
    for ii in psi_rows:
        PSI[:, ii] = X[ii * nr:(ii + 1) * nr]
    for ii in ts_rows:
        T[:, ii] = X[N + ii * nr:N + (ii + 1) * nr]
        C[:, ii] = X[2 * N + ii * nr:2 * N + (ii + 1) * nr]
    return PSI, T, C


# ------------------------------------------------------------ theta operators
def J_theta_RT(g, nr, N_fm, symmetric):
    """(1/sin)d_theta(sin psi): sine rows -> cosine rows (Matrix_Operators.py:437-472)."""
    p = _reg.theta_plan(N_fm, nr, symmetric)
    return _host(p.theta_ops(_dev(_psi_state(g, N_fm * nr))))[0, 0]


def DT0_theta(g, dT0, N_fm, nr, symmetric):
    """r^2 J(psi, T0) = dT0 * J_theta_RT(psi) (Matrix_Operators.py:132-189)."""
    p = _reg.theta_plan(N_fm, nr, symmetric, dT0=dT0)
    return _host(p.theta_ops(_dev(_psi_state(g, N_fm * nr))))[0, 1]


def A2_SINE(g, D, R, N_fm, nr, symmetric):
    """E^2 psi in the sine basis (Matrix_Operators.py:193-245)."""
    p = _reg.geometry_plan(D, R, N_fm, symmetric)
    return _host(p.theta_ops(_dev(_psi_state(g, N_fm * nr))))[0, 2]


def A2_SINE_R2(g, N_fm, nr, D, R, symmetric):
    """(1/r^2) E^2 psi in the sine basis (Matrix_Operators.py:476-526)."""
    p = _reg.geometry_plan(D, R, N_fm, symmetric)
    return _host(p.theta_ops(_dev(_psi_state(g, N_fm * nr))))[0, 3]


# ------------------------------------------------------- nonlinear advection
def Derivatives(X_hat, JPSI, OMEGA, Dr, N_fm, nr, symmetric):
    """The nine (nr, N_fm) slabs JT_psi, kDpsi, komega, DT, DC, omega, Dpsi,
    kT, kC (Matrix_Operators.py:631-740)."""
    D, R = _reg.geometry_of_dr(Dr)
    p = _reg.geometry_plan(D, R, N_fm, symmetric)
    s = _host(p.derivatives(_dev(X_hat)))[0]
    return tuple(np.array(s[q]) for q in range(9))


def NLIN_FX(X_hat, D, R, N_fm, nr, symmetric):
    """Nonlinear advection terms F(X) (Matrix_Operators.py:743-804)."""
    if N_fm % 2 != 0:
        raise ValueError("The number of Fourier modes is not even %d" % N_fm)
    p = _reg.geometry_plan(D, R, N_fm, symmetric)
    return _host(p.nlin_fx(_dev(X_hat)))[0]


def NLIN_DFX(dv_hat, X_hat, D, R, N_fm, nr, symmetric):
    """F(X, dv) + F(dv, X) (Matrix_Operators.py:807-898)."""
    if N_fm % 2 != 0:
        raise ValueError("The number of Fourier modes is not even %d" % N_fm)
    p = _reg.geometry_plan(D, R, N_fm, symmetric)
    return _host(p.nlin_dfx(_dev(dv_hat), _dev(X_hat)))[0]


# ------------------------------------------------------ implicit back solves
def NAB2_BSub_TSTEP_V2(g, L_inv, N_fm, nr, dt, symmetric):
    """Solve r^2(1 - dt Lap) f = g with the per-mode inverses L_inv
    (Matrix_Operators.py:1034-1086)."""
    p = _reg.nab2_plan(L_inv, N_fm, nr, dt, symmetric)
    return _host(p.bsub(_dev(g), 1))[0]


def A4_BSub_TSTEP_V2(g, L_inv, D2, IR4, IR2, N_fm, nr, dt, symmetric):
    """Solve (A^2 - dt A^4) psi = g (dt = Pr dt as the caller passes it)
    (Matrix_Operators.py:1116-1194)."""
    p = _reg.a4_plan(L_inv, D2, IR4, IR2, N_fm, nr, dt, symmetric)
    return _host(p.bsub(_dev(g), 0))[0]


# ------------------------------------------------------ resolution changes
def INTERP_RADIAL(N_n, N_o, X_o, d):
    """Move a state between radial resolutions by a degree-len(R_o) polynomial
    fit through the old points with zero boundary values
    (Matrix_Operators.py:901-941; host NumPy, as a file-load step)."""
    import warnings
    if N_n == N_o:
        return X_o
    print('Interpolated in r from %d to %d' % (N_o, N_n), '\n')
    _, R_n = _ops.cheb_radial(N_n, d)
    _, R_o = _ops.cheb_radial(N_o, d)
    nr_n, nr_o = len(R_n[1:-1]), len(R_o[1:-1])
    N_fm = len(X_o) // (3 * nr_o)
    X_n = np.zeros(3 * nr_n * N_fm)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")   # the fit is rank deficient by design
        for f in range(3):
            for k in range(N_fm):
                io = f * N_fm * nr_o + nr_o * k
                c = np.polyfit(R_o, np.hstack(([0.0], X_o[io:io + nr_o], [0.0])), len(R_o))
                i_n = f * N_fm * nr_n + nr_n * k
                X_n[i_n:i_n + nr_n] = np.polyval(c, R_n[1:-1])
    return X_n


def INTERP_THETAS(N_fm_n, N_fm_o, X_o):
    """Move a state between theta resolutions through the grid: inverse
    transforms at the old size, forward transforms truncated to the new one
    (Matrix_Operators.py:944-1011; transforms on the GPU)."""
    from .Transforms import DCT, DST, IDCT, IDST
    print('Interpolated in theta from %d to %d' % (N_fm_o, N_fm_n), '\n')
    if N_fm_n == N_fm_o:
        return X_o
    nr = len(X_o) // (3 * N_fm_o)
    width = max(N_fm_n, N_fm_o)
    fields = [np.zeros((nr, width)) for _ in range(3)]
    for f in range(3):
        for k in range(N_fm_o):
            ind = f * N_fm_o * nr + k * nr
            fields[f][:, k] = X_o[ind:ind + nr]
    psi_hat = DST(IDST(fields[0]), n=N_fm_n)
    t_hat = DCT(IDCT(fields[1]), n=N_fm_n)
    s_hat = DCT(IDCT(fields[2]), n=N_fm_n)
    XX = np.zeros(3 * nr * N_fm_n)
    for f, h in enumerate((psi_hat, t_hat, s_hat)):
        for k in range(N_fm_n):
            ind = f * N_fm_n * nr + k * nr
            XX[ind:ind + nr] = h[:, k]
    return XX
