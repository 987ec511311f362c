"""Host-side construction of the radial operators (set-up, float64, NumPy).

The operators are built with the same NumPy operations, in the same order, as
the reference's Build_Matrix_Operators (Main.py:179-225), so that the
pre-inverted blocks uploaded to the GPU are the reference's own matrices:
  cheb_radial      Matrix_Operators.py:10-28
  Nabla2, Nabla4   Matrix_Operators.py:31-76
  A4_TSTEP_MATS    Matrix_Operators.py:1089-1112
  NAB2_TSTEP_MATS  Matrix_Operators.py:1014-1030
This is one-time set-up (K matrix inversions of size nr); every per-step
operation runs in libsddc.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def cheb_radial(N, d):
    """Chebyshev-Gauss-Lobatto points on [1/d, (1+d)/d] and Trefethen's
    differentiation matrix with the negative-row-sum diagonal
    (Matrix_Operators.py:10-28)."""
    r_i = 1.0 / d
    r_o = (1.0 + d) / d
    if N == 0:
        return 0.0, 1.0
    n = np.arange(0, N + 1)
    x = np.cos(np.pi * n / N).reshape(N + 1, 1)
    x = 0.5 * (r_i + r_o) + 0.5 * (r_i - r_o) * x
    c = (np.hstack(([2.0], np.ones(N - 1), [2.0])) * (-1) ** n).reshape(N + 1, 1)
    xx = np.tile(x, (1, N + 1))
    D = np.dot(c, 1.0 / c.T) / (xx - xx.T + np.eye(N + 1))
    D -= np.diag(np.sum(D.T, axis=0))
    return D, x.reshape(N + 1)


def nabla2(D, r):
    """Interior block of r^2 d2/dr2 + 2 r d/dr (Matrix_Operators.py:31-45)."""
    A = np.diag(r[:] ** 2) @ (D @ D) + np.diag(2.0 * r[:]) @ D
    return A[1:-1, 1:-1]


def nabla4(D, r):
    """Clamped fourth derivative (Matrix_Operators.py:48-76)."""
    one = np.ones(len(r))
    b = -(r[0] + r[-1])
    c = r[0] * r[-1]
    q = (r ** 2) + b * r + c * one
    with np.errstate(divide="ignore"):
        S = np.diag(1.0 / q)
    S[0, 0] = 0.0
    S[-1, -1] = 0.0
    D2 = D @ D
    D3 = D @ D2
    D4 = D2 @ D2
    L4 = np.diag(q) @ D4 + 4.0 * np.diag(2.0 * r + b * one) @ D3 + 12.0 * D2
    return (L4 @ S)[1:-1, 1:-1]


def base_state_coeffs(d):
    """T_0 = -A_T/r + B_T (Main.py:23-38)."""
    R_1 = 1.0 / d
    R_2 = (1.0 + d) / d
    return (R_1 * R_2) / (R_1 - R_2), R_1 / (R_1 - R_2)


def a4_inverses(dt, N_fm, nr, D, R):
    """Per-mode inverses of (A2 + bj IR2) - dt (D4 + bj (D2' + bj IR4)),
    j = N_fm - jj (Matrix_Operators.py:1089-1112); dt is Pr*dt."""
    IR = np.diag(1.0 / R)
    IR2 = IR @ IR
    D_sq = D @ D
    D4 = nabla4(D, R)
    D2 = (IR2 @ (2 * D_sq - 4 * (IR @ D) + 6 * IR2))[1:-1, 1:-1]
    A2 = D_sq[1:-1, 1:-1]
    IR2 = IR2[1:-1, 1:-1]
    IR4 = IR2 @ IR2
    out = np.empty((N_fm, nr, nr))
    for jj in range(N_fm):
        j = N_fm - jj
        bj = -j * (j + 1)
        L1 = D2 + bj * IR4
        out[jj] = np.linalg.inv((A2 + bj * IR2) - dt * (D4 + bj * L1))
    return out


def nab2_inverses(dt, N_fm, nr, D, R):
    """Per-mode inverses of r^2 - dt (r^2 Lap - j(j+1)), j = N_fm-1-jj
    (Matrix_Operators.py:1014-1030)."""
    I = np.eye(nr)
    R2 = np.diag(R[1:-1] ** 2)
    R2_Nab2 = nabla2(D, R)
    out = np.empty((N_fm, nr, nr))
    for jj in range(N_fm):
        j = N_fm - (jj + 1)
        bj = -j * (j + 1)
        out[jj] = np.linalg.inv(R2 - dt * (R2_Nab2 + bj * I))
    return out


def nab2_modes(dt, nr, D, R):
    """Eigenbasis of P = r^2 - dt r^2 Lap, the mode-independent part of the
    NAB2 matrices: the block of mode j is P + dt j(j+1) I
    (Matrix_Operators.py:1014-1030), so Linv_j = V diag(1/(lam + dt j(j+1))) V^-1.
    Returns one array: lam (nr), then V and V^-1 (nr x nr each, row-major), or
    None when P has complex eigenvalues (the solve then uses the blocks)."""
    P = np.diag(R[1:-1] ** 2) - dt * nabla2(D, R)
    lam, V = np.linalg.eig(P)
    if np.iscomplexobj(lam) and np.abs(lam.imag).max() > 1e-12 * np.abs(lam).max():
        return None
    lam, V = lam.real, np.real(V)
    return np.ascontiguousarray(np.concatenate([lam, V.ravel(), np.linalg.inv(V).ravel()]))


@dataclass
class OperatorSet:
    N_fm: int
    N_r: int
    d: float
    dt: float
    Pr: float
    Tau: float
    D: np.ndarray
    R: np.ndarray
    Dr: np.ndarray
    D2s: np.ndarray
    D2r2: np.ndarray
    IR2v: np.ndarray
    IR4v: np.ndarray
    DT0: np.ndarray
    Rsq: np.ndarray
    GR: np.ndarray
    D2_A4: np.ndarray
    IR2_A4: np.ndarray
    IR4_A4: np.ndarray
    Linv_A4: np.ndarray
    Linv_T: np.ndarray
    Linv_S: np.ndarray
    modes_T: np.ndarray = None     # nab2_modes(dt): the T solve in its eigenbasis
    modes_S: np.ndarray = None     # nab2_modes(Tau*dt)

    @property
    def nr(self):
        return self.N_r - 1


def radial_pieces(D, R, d):
    """Everything that depends on (D, R) only."""
    ri = R[1:-1]
    A_T, _ = base_state_coeffs(d)
    IR = np.diag(1.0 / R)
    IR2f = IR @ IR
    D_sq = D @ D
    D2_A4 = (IR2f @ (2 * D_sq - 4 * (IR @ D) + 6 * IR2f))[1:-1, 1:-1]
    IR2 = IR2f[1:-1, 1:-1]
    IR4 = IR2 @ IR2
    R_1 = 1.0 / d
    return dict(
        Dr=np.ascontiguousarray(D[1:-1, 1:-1]),
        D2s=np.ascontiguousarray((D @ D)[1:-1, 1:-1]),                      # A2_SINE :217
        D2r2=np.ascontiguousarray((np.diag(1.0 / R ** 2) @ (D @ D))[1:-1, 1:-1]),  # :498
        IR2v=1.0 / (ri ** 2),
        IR4v=1.0 / (ri ** 4),
        DT0=A_T / (ri ** 2),                                                # Main.py:201-202
        Rsq=ri ** 2,
        GR=(R_1 ** 2) / (ri ** 2),                                          # kGR_RT :113
        D2_A4=np.ascontiguousarray(D2_A4),
        IR2_A4=np.ascontiguousarray(np.diag(IR2).copy()),
        IR4_A4=np.ascontiguousarray(np.diag(IR4).copy()),
    )


def build(N_fm, N_r, d, dt, Pr, Tau):
    """Operator set of Build_Matrix_Operators(N_fm, N_r, d, dt, Pr, Tau)."""
    if N_fm % 2 != 0:
        raise ValueError("The number of Fourier modes is not even %d" % N_fm)
    D, R = cheb_radial(N_r, d)
    nr = N_r - 1
    pieces = radial_pieces(D, R, d)
    return OperatorSet(
        N_fm=N_fm, N_r=N_r, d=d, dt=dt, Pr=Pr, Tau=Tau, D=np.ascontiguousarray(D), R=R,
        Linv_A4=a4_inverses(Pr * dt, N_fm, nr, D, R),
        Linv_T=nab2_inverses(dt, N_fm, nr, D, R),
        Linv_S=nab2_inverses(Tau * dt, N_fm, nr, D, R),
        modes_T=nab2_modes(dt, nr, D, R),
        modes_S=nab2_modes(Tau * dt, nr, D, R),
        **pieces,
    )


def from_matrices(D, R, N_fm, d=None, dt=0.0, Pr=1.0, Tau=1.0, Linv_A4=None, Linv_T=None,
                  Linv_S=None, D2_A4=None, IR2_A4=None, IR4_A4=None):
    """Operator set around caller-supplied matrices, for the per-function
    drop-ins (which receive D, R, L_inv ... rather than parameters).  Blocks
    that are not supplied are zero."""
    D = np.ascontiguousarray(np.asarray(D, dtype=np.float64))
    R = np.ascontiguousarray(np.asarray(R, dtype=np.float64))
    N_r = len(R) - 1
    nr = N_r - 1
    if d is None:
        d = 1.0 / R[0]
    pieces = radial_pieces(D, R, d)
    z = np.zeros((N_fm, nr, nr))
    if D2_A4 is not None:
        pieces["D2_A4"] = np.ascontiguousarray(np.asarray(D2_A4, dtype=np.float64))
    if IR2_A4 is not None:
        pieces["IR2_A4"] = np.ascontiguousarray(np.diag(np.asarray(IR2_A4, dtype=np.float64)).copy())
    if IR4_A4 is not None:
        pieces["IR4_A4"] = np.ascontiguousarray(np.diag(np.asarray(IR4_A4, dtype=np.float64)).copy())

    def stack(L):
        if L is None:
            return z
        return np.ascontiguousarray(np.array([np.asarray(m, dtype=np.float64) for m in L]))

    return OperatorSet(N_fm=N_fm, N_r=N_r, d=d, dt=dt, Pr=Pr, Tau=Tau, D=D, R=R,
                       Linv_A4=stack(Linv_A4), Linv_T=stack(Linv_T), Linv_S=stack(Linv_S),
                       **pieces)
