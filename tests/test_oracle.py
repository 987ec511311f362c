"""Pins of the CPU oracle (oracle/sddc_oracle.py) -- no GPU needed.

Each check is chosen so that a plausible mistake (dropped term, wrong sign,
index or parity) fails it:
  * the reference's own transform known-answer tests (Transforms.py:132-374);
  * brute-force O(K*M) sums of the transform definitions;
  * Chebyshev differentiation exact on polynomials;
  * Legendre eigenfunctions, which reduce the theta-coupled recurrences
    (J_theta_RT, A2_SINE, A2_SINE_R2, both back substitutions) to textbook
    radial operators;
  * golden vectors written from the reference (tests/golden/make_golden.py);
  * algebraic invariants (bilinearity, directional derivative, symmetry);
  * the reference's linear growth-rate target (Tests/Run_Tests.py:55-65).
"""
import os

import numpy as np
import pytest
from numpy.polynomial import chebyshev as Ch
from numpy.polynomial import legendre as Lg

from oracle import sddc_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = dict(np.load(os.path.join(HERE, "golden", "golden_c1.npz")))


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _c1():
    g = GOLD
    return (int(g["N_r"]), int(g["N_fm"]), float(g["d"]), float(g["Pr"]), float(g["Tau"]),
            float(g["Ra"]), float(g["Ra_s"]), float(g["dt"]))


# ----------------------------------------------------------------- transforms
N_KAT = 256


def _eq12(a, b):
    assert np.array_equal(np.round(a, 12), np.round(b, 12))


@pytest.mark.parametrize("k", [0, 1, 2])
def test_transform_kat_roundtrip(k):
    """Transforms.py:132-196 single-mode round trips."""
    N = N_KAT
    x = O.grid(N)
    f_hat = np.zeros(N)
    f_hat[k] = 1
    _eq12(O.dct(O.idct(f_hat)), f_hat)
    _eq12(O.idct(O.dct(np.cos(k * x))), np.cos(k * x))
    g_hat = np.zeros(N)
    g_hat[k + 1] = 1
    _eq12(O.dst(O.idst(g_hat)), g_hat)
    _eq12(O.idst(O.dst(np.sin((k + 1) * x))), np.sin((k + 1) * x))


@pytest.mark.parametrize("k", [0, 1, 2])
def test_transform_kat_dealias(k):
    """Transforms.py:199-261: 3/2-padded synthesis then truncated analysis."""
    N = N_KAT
    f_hat = np.zeros(N)
    f_hat[k] = 2.1
    _eq12(O.dct(O.idct(f_hat, n=(3 * N) // 2), n=N), f_hat)
    g_hat = np.zeros(N)
    g_hat[k + 1] = 3.3
    _eq12(O.dst(O.idst(g_hat, n=(3 * N) // 2), n=N), g_hat)


def test_transform_kat_products():
    """Transforms.py:264-374: products of modes land on exact coefficients."""
    N = N_KAT
    x = O.grid(N)
    g_hat = np.zeros(N)
    g_hat[2] += 0.5
    g_hat[1] += -0.5 + 0.5
    g_hat[3] += 0.5 + 0.5
    g = np.sin(x) * np.cos(x) + np.sin(x) * np.cos(2 * x) + np.sin(2 * x) * np.cos(x)
    _eq12(O.dst(g, n=N), g_hat)
    _eq12(O.idst(g_hat, n=N), g)
    f_hat = np.zeros(N)
    f_hat[[0, 1, 2, 3]] = [1 + 0.5, 1 + 0.5, 1 - 0.5, 1 - 0.5]
    f = (2 * np.cos(x) * np.cos(2 * x) + 2 * np.cos(x) ** 2 + np.sin(x) * np.sin(2 * x)
         + np.sin(x) ** 2)
    _eq12(O.dct(f, n=N), f_hat)
    _eq12(O.idct(f_hat, n=N), f)


@pytest.mark.parametrize("K,M", [(8, 12), (10, 15), (16, 24), (6, 6)])
def test_transform_brute_force(K, M):
    rng = np.random.default_rng(K * M)
    th = np.pi * (2 * np.arange(M) + 1) / (2 * M)
    a = rng.standard_normal(K)
    x = rng.standard_normal(M)
    k = np.arange(K)
    assert rel(O.idct(a, M), np.cos(np.outer(th, k)) @ a) < 1e-14
    b = a.copy()
    b[0] = 0.0
    assert rel(O.idst(a, M), np.sin(np.outer(th, k)) @ b) < 1e-14
    m = np.arange(M)
    c = (2.0 / M) * (np.cos(np.outer(m, th)) @ x)
    c[0] *= 0.5
    assert rel(O.dct(x, K), c[:K]) < 1e-14
    s = (2.0 / M) * (np.sin(np.outer(m, th)) @ x)
    assert rel(O.dst(x, K), s[:K]) < 1e-14


# ------------------------------------------------------------ radial operators
@pytest.mark.parametrize("N", [8, 20, 40])
def test_cheb_exact_on_polynomials(N):
    D, R = O.cheb_radial(N, 0.35)
    assert R[0] < R[-1]
    for p in range(N + 1):
        f = (R - R.mean()) ** p
        df = p * (R - R.mean()) ** max(p - 1, 0) if p else 0 * R
        assert np.abs(D @ f - df).max() <= 1e-9 * max(1.0, np.abs(df).max())


# ------------------------------------------------- Legendre eigenfunctions
def _cos_coefs(l, K):
    c = Ch.poly2cheb(Lg.leg2poly(np.eye(l + 1)[l]))
    a = np.zeros(K)
    a[:len(c)] = c
    return a


def _sine_rows(l, K):
    a = _cos_coefs(l, K)
    s = np.zeros(K)
    for m in range(1, K):
        s[m - 1] = m * a[m]
    return s


@pytest.mark.parametrize("l", [1, 2, 3, 4, 7, 10])
def test_theta_recurrences_legendre(l):
    """psi = f(r) (-dP_l/dtheta), T = f(r) P_l(cos theta) are eigenfunctions:
    J_theta(psi) = l(l+1) T-shape; E^2 psi = (D2 f - l(l+1) f/r^2) psi-shape;
    the two back substitutions invert r^2(1 - dt Lap) and A^2 - dt A^4 on them."""
    K, Nr, d, dt, Pr = 16, 12, 0.5, 0.01, 1.0
    nr = Nr - 1
    o = O.build_operators(K, Nr, d, dt, Pr, 1.0)
    R, D = o.R, o.D
    ri = R[1:-1]
    L = l * (l + 1.0)
    S, C = _sine_rows(l, K), _cos_coefs(l, K)
    f = np.sin(np.pi * (ri - ri[0] + 0.3)) * (ri - R[0]) * (R[-1] - ri)
    psi = np.outer(S, f).ravel()
    scale = np.abs(S).max() * np.abs(f).max()
    tol = 1e-13 * (1 + l)
    J = O.j_theta(psi, K, nr).reshape(K, nr)
    assert np.abs(J - L * np.outer(C, f)).max() <= tol * L * scale
    Dsq = D @ D
    A2 = O.a2_sine(psi, o).reshape(K, nr)
    ref = np.outer(S, Dsq[1:-1, 1:-1] @ f - L * f / ri ** 2)
    assert np.abs(A2 - ref).max() <= tol * np.abs(ref).max()
    A2R2 = O.a2_sine_r2(psi, o).reshape(K, nr)
    ref = np.outer(S, (np.diag(1 / R ** 2) @ Dsq)[1:-1, 1:-1] @ f - L * f / ri ** 4)
    assert np.abs(A2R2 - ref).max() <= tol * np.abs(ref).max()
    T = np.outer(C, f).ravel()
    opT = np.diag(ri ** 2) - dt * (O.nabla2(D, R) - L * np.eye(nr))
    g = np.outer(C, opT @ f).ravel()
    assert np.abs(O.nab2_bsub(g, o.Linv_T, K, nr, dt) - T).max() <= 1e-12 * np.abs(T).max()
    IR = np.diag(1 / R)
    IR2f = IR @ IR
    D2p = (IR2f @ (2 * Dsq - 4 * (IR @ D) + 6 * IR2f))[1:-1, 1:-1]
    IR2 = IR2f[1:-1, 1:-1]
    IR4 = IR2 @ IR2
    opA = (Dsq[1:-1, 1:-1] - L * IR2) - Pr * dt * (O.nabla4(D, R) - L * (D2p - L * IR4))
    gA = np.outer(S, opA @ f).ravel()
    out = O.a4_bsub(gA, o.Linv_A4, o.D2_A4, o.IR4_A4, o.IR2_A4, K, nr, Pr * dt)
    assert np.abs(out - psi).max() <= 1e-11 * np.abs(psi).max()


def test_nusselt_closed_form():
    """T = f(r) cos(k theta): Nu-1 at R_1 = (R_1^2/A_T) f'(R_1) / (1 - k^2)
    for even k (Main.py:52-68)."""
    K, Nr, d = 8, 16, 0.35
    nr = Nr - 1
    D, R = O.cheb_radial(Nr, d)
    A_T, _ = O.base_state_coeffs(d)
    f = (R - R[0]) * (R[-1] - R) * (1 + R)
    fp0 = (R[-1] - R[0]) * (1 + R[0])
    for k in (0, 2, 4):
        T = np.zeros(K * nr)
        T[k * nr:(k + 1) * nr] = f[1:-1]
        nu = O.nusselt(T, d, R, D, K, nr, check=False)
        assert abs(nu - (R[0] ** 2 / A_T) * fp0 / (1 - k * k)) <= 1e-10 * abs(nu)


# --------------------------------------------------------------- golden vectors
def test_golden_slabs_and_nonlinear():
    N_r, K, d, Pr, Tau, Ra, Ras, dt = _c1()
    o = O.build_operators(K, N_r, d, dt, Pr, Tau)
    s = O.spectral_slabs(GOLD["X_big"], o)
    for k in s:
        assert rel(s[k], GOLD["slab_" + k]) < 1e-14, k
    for sym in (False, True):
        tag = "sym" if sym else "full"
        assert rel(O.nlin_fx(GOLD["X_big"], o, sym), GOLD["nlin_fx_big_" + tag]) < 1e-14
        assert rel(O.nlin_fx(GOLD["X_small"], o, sym), GOLD["nlin_fx_small_" + tag]) < 1e-14
        assert rel(O.nlin_dfx(GOLD["dv"], GOLD["X_big"], o, sym), GOLD["nlin_dfx_" + tag]) < 1e-14
        ke = O.kinetic_energy(GOLD["X_big"], o.R, o.D, K, N_r - 1, sym)
        assert abs(ke / GOLD["ke_big_" + tag] - 1) < 1e-13


@pytest.mark.parametrize("sym", [False, True])
def test_golden_trajectory(sym):
    """20 steps of the reference's _Time_Step with its per-step diagnostics."""
    N_r, K, d, Pr, Tau, Ra, Ras, dt = _c1()
    tag = "sym" if sym else "full"
    X, h = O.time_step(GOLD["X_small"], Ra, Ras, Tau, Pr, d, K, N_r, sym, 20, dt)
    assert rel(X, GOLD["traj20_X_" + tag]) < 1e-13
    for key in ("Norm", "KE", "Nu_T", "Nu_S"):
        assert np.max(np.abs(np.array(h[key]) / GOLD["traj20_%s_%s" % (key, tag)] - 1)) < 1e-12


def test_golden_trajectory_large_and_linear():
    N_r, K, d, Pr, Tau, Ra, Ras, dt = _c1()
    X, h = O.time_step(GOLD["X_big"], Ra, Ras, Tau, Pr, d, K, N_r, False, 10, dt)
    assert rel(X, GOLD["traj10big_X_full"]) < 1e-12
    X, _ = O.time_step(GOLD["X_small"], Ra, Ras, Tau, Pr, d, K, N_r, False, 10, dt, linear=True)
    assert rel(X, GOLD["traj10lin_X_full"]) < 1e-13


# ------------------------------------------------------------------ invariants
def test_nonlinear_invariants():
    N_r, K, d, Pr, Tau, Ra, Ras, dt = _c1()
    o = O.build_operators(K, N_r, d, dt, Pr, Tau)
    rng = np.random.default_rng(3)
    X = rng.standard_normal(3 * K * (N_r - 1))
    V = rng.standard_normal(3 * K * (N_r - 1))
    F = O.nlin_fx(X, o)
    assert np.abs(O.nlin_fx(0 * X, o)).max() == 0.0
    assert rel(O.nlin_fx(2 * X, o), 4 * F) < 1e-13
    # F quadratic => F(X+V) = F(X) + F(V) + DF(X) V
    lhs = O.nlin_fx(X + V, o) - F - O.nlin_fx(V, o)
    assert rel(O.nlin_dfx(V, X, o), lhs) < 1e-11
    assert rel(O.nlin_dfx(V, X, o), O.nlin_dfx(X, V, o)) < 1e-13


def test_symmetric_is_masked_full_step():
    N_r, K, d, Pr, Tau, Ra, Ras, dt = _c1()
    o = O.build_operators(K, N_r, d, dt, Pr, Tau)
    m = O.eq_sym_mask(K, N_r - 1)
    X = GOLD["X_small"]
    a = O.step(X, o, Ra, Ras, symmetric=True)
    b = m * O.step(m * X, o, Ra, Ras)
    assert rel(a, b) < 1e-15


def test_linear_growth_rate_target():
    """Tests/Run_Tests.py:37-65: even case d=2, Ra=7268.365, Ra_s=500, Tau=1,
    Pr=1, N_r=20, N_fm=10, symmetric, linear; target lambda = 0.0018195.  The
    scheme is first order in dt, so the Richardson extrapolation of dt=5e-3
    and 2.5e-3 must land on the target."""
    K, Nr = 10, 20
    nr = Nr - 1
    lam = []
    for dt in (5e-3, 2.5e-3):
        o = O.build_operators(K, Nr, 2.0, dt, 1.0, 1.0)
        X = np.random.default_rng(1).random(3 * K * nr)
        X = 1e-3 * X / np.linalg.norm(X)
        m = O.eq_sym_mask(K, nr)
        n = int(round(40 / dt))
        norms = np.empty(n)
        for it in range(n):
            Xn = O.step(X, o, 7268.365, 500.0, linear=True, symmetric=True)
            norms[it] = np.linalg.norm(Xn)
            X = m * Xn
        t = dt * np.arange(n)
        lam.append(np.polyfit(t[-1000:], np.log(norms[-1000:]), 1)[0])
    extrap = 2 * lam[1] - lam[0]
    assert abs(extrap / 0.0018195 - 1) < 0.02, lam


@pytest.mark.parametrize("sym", [False, True])
def test_step_linearized_is_the_central_difference_of_the_step(sym):
    """The nonlinear term is quadratic, so the step is a quadratic map of X
    and (step(X + h v) - step(X - h v)) / (2h) equals the linearised step
    (PDFX + v, Main.py:498-521) exactly, up to rounding, for any h."""
    K, Nr = 16, 12
    o = O.build_operators(K, Nr, 0.35, 5e-3, 1.0, 1 / 15.0)
    rng = np.random.default_rng(17)
    n3 = 3 * K * (Nr - 1)
    X = 0.3 * rng.standard_normal(n3)
    v = rng.standard_normal(n3)
    lin = O.step_linearized(v, X, o, 3000.0, 500.0, symmetric=sym)
    for h in (1.0, 0.1, 1e-3):
        cd = (O.step(X + h * v, o, 3000.0, 500.0, symmetric=sym)
              - O.step(X - h * v, o, 3000.0, 500.0, symmetric=sym)) / (2 * h)
        assert np.linalg.norm(cd - lin) / np.linalg.norm(lin) < 1e-11 / h, h
    # a plausible slip (the factor 2 of F(X, v) + F(v, X) lost) fails it
    half = 0.5 * (lin + O.step_linearized(v, 0 * X, o, 3000.0, 500.0, symmetric=sym))
    cd = (O.step(X + v, o, 3000.0, 500.0, symmetric=sym)
          - O.step(X - v, o, 3000.0, 500.0, symmetric=sym)) / 2
    assert np.linalg.norm(cd - half) / np.linalg.norm(lin) > 1e-6
