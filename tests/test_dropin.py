"""The drop-in modules (Transforms, Matrix_Operators, Main under the
reference's names) against the oracle and the reference's golden vectors.

CPU tests cover the host-side parts (layout helpers, symmetry mask, sparse
operator builders, module registration); GPU tests cover every function that
computes on the device.  Tolerances as in test_gpu_parity.py: 1e-14 for
transforms, theta operators and solves, 1e-13 for nonlinear terms, 1e-12 per
step, 1e-12 on diagnostics of a golden trajectory.
"""
import os
import sys

import numpy as np
import pytest

from oracle import sddc_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = dict(np.load(os.path.join(HERE, "golden", "golden_c1.npz")))
C1 = dict(N_r=20, N_fm=64, d=0.35, Pr=1.0, Tau=1.0 / 15.0, Ra=9000.0, Ra_s=500.0, dt=5e-3)
K, NR = C1["N_fm"], C1["N_r"] - 1
N = K * NR


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def mods():
    from spectraldoublediffusiveconvection_amd import dropin
    m = dropin.install()
    yield m
    dropin.uninstall()


@pytest.fixture(scope="module")
def o():
    return O.build_operators(C1["N_fm"], C1["N_r"], C1["d"], C1["dt"], C1["Pr"], C1["Tau"])


# ------------------------------------------------------------------ CPU tests
def test_install_registers_reference_names(mods):
    import Main
    import Matrix_Operators
    import Transforms
    assert Main is mods["Main"] and Transforms is mods["Transforms"]
    assert Matrix_Operators is mods["Matrix_Operators"]
    for name in ("_Time_Step", "_Newton", "Kinetic_Energy", "Nusselt", "Eq_SYM",
                 "Build_Matrix_Operators", "Base_State_Coeffs"):
        assert callable(getattr(Main, name))
    for name in ("cheb_radial", "Nabla2", "Nabla4", "R2", "kGR_RT", "DT0_theta", "A2_SINE",
                 "J_theta_RT", "A2_SINE_R2", "Vecs_to_X", "X_to_Vecs", "Derivatives", "NLIN_FX",
                 "NLIN_DFX", "NAB2_TSTEP_MATS", "A4_TSTEP_MATS", "NAB2_BSub_TSTEP_V2",
                 "A4_BSub_TSTEP_V2", "INTERP_RADIAL", "INTERP_THETAS"):
        assert callable(getattr(Matrix_Operators, name))
    for name in ("grid", "DCT", "IDCT", "DST", "IDST", "sinusoid_to_IDCT", "DCT_to_sinusoid",
                 "sinusoid_to_IDST", "DST_to_sinusoid"):
        assert callable(getattr(Transforms, name))


def test_host_parts_match_oracle(mods, o):
    Main, MO, TR = mods["Main"], mods["Matrix_Operators"], mods["Transforms"]
    assert np.array_equal(TR.grid(96), O.grid(96))
    D, R = MO.cheb_radial(C1["N_r"], C1["d"])
    assert np.array_equal(D, o.D) and np.array_equal(R, o.R)
    assert np.array_equal(Main.Eq_SYM(np.zeros(3 * N), R), O.eq_sym_mask(K, NR))
    ops = Main.Build_Matrix_Operators(K, C1["N_r"], C1["d"], C1["dt"], C1["Pr"], C1["Tau"])
    D_, R_, Rsq, DT0, gr_K, LA4, LT, LS, args_A4, args_FX = ops
    assert np.array_equal(np.asarray(LA4), o.Linv_A4) and np.array_equal(np.asarray(LT), o.Linv_T)
    assert np.array_equal(np.asarray(LS), o.Linv_S) and np.array_equal(DT0, o.DT0)
    assert np.array_equal(args_A4[0], o.D2_A4)
    v = np.random.default_rng(0).standard_normal(N)
    assert np.allclose(Rsq.dot(v), np.tile(o.Rsq, K) * v, rtol=0, atol=0)
    ref = np.zeros((K, NR))
    ref[:K - 1] = o.grk[:K - 1] * v.reshape(K, NR)[1:]
    assert np.array_equal(gr_K.dot(v), ref.reshape(-1))
    # layout helpers: round trip, and the symmetric variants keep only the kept rows
    X = np.random.default_rng(1).standard_normal(3 * N)
    for sym in (False, True):
        P, T, C = MO.X_to_Vecs(X, K, NR, sym)
        back = MO.Vecs_to_X(P, T, C, K, NR, sym)
        expect = X * (O.eq_sym_mask(K, NR) if sym else 1.0)
        assert np.array_equal(back, expect)


def test_scaling_helpers(mods):
    TR = mods["Transforms"]
    a = np.random.default_rng(2).standard_normal((3, 10))
    s = TR.sinusoid_to_IDCT(a)
    assert np.array_equal(s[:, 0], a[:, 0]) and np.array_equal(s[:, 1:], 0.5 * a[:, 1:])
    s = TR.sinusoid_to_IDST(a)
    assert np.array_equal(s[:, :-1], 0.5 * a[:, 1:]) and not s[:, -1].any()
    s = TR.DST_to_sinusoid(a)
    assert np.array_equal(s[:, 1:], a[:, :-1] * (1 / 10)) and not s[:, 0].any()
    s = TR.DCT_to_sinusoid(a)
    assert np.array_equal(s[:, 0], a[:, 0] * (1 / 10) * 0.5)
    with pytest.raises(ValueError):
        TR.DCT(a, axis=0)


def test_interp_radial_exact_on_polynomials(mods):
    """A field that is a polynomial in r of degree <= N_o vanishing at both
    walls is carried exactly to the new grid."""
    MO = mods["Matrix_Operators"]
    d, N_o, N_n, Kf = 0.35, 12, 16, 4
    _, R_o = O.cheb_radial(N_o, d)
    _, R_n = O.cheb_radial(N_n, d)
    f = lambda r: (r - R_o[0]) * (r - R_o[-1]) * (1 + 0.3 * r - 0.05 * r ** 2)
    X_o = np.concatenate([np.tile(f(R_o[1:-1]) * (k + 1), 1) for k in range(3 * Kf)])
    X_n = MO.INTERP_RADIAL(N_n, N_o, X_o, d)
    expect = np.concatenate([f(R_n[1:-1]) * (k + 1) for k in range(3 * Kf)])
    assert rel(X_n, expect) < 1e-8
    assert MO.INTERP_RADIAL(N_o, N_o, X_o, d) is X_o


# ------------------------------------------------------------------ GPU tests
@pytest.mark.gpu
def test_transforms_dropin(mods):
    TR = mods["Transforms"]
    rng = np.random.default_rng(3)
    a = rng.standard_normal((5, 64))
    x = rng.standard_normal((5, 96))
    assert rel(TR.IDCT(a, n=96), O.idct(a, 96)) < 1e-14
    assert rel(TR.IDST(a, n=96), O.idst(a, 96)) < 1e-14
    assert rel(TR.DCT(x, n=64), O.dct(x)[:, :64]) < 1e-14
    assert rel(TR.DST(x, n=64), O.dst(x)[:, :64]) < 1e-14
    assert rel(TR.IDCT(a[0]), O.idct(a[0])) < 1e-14          # 1-D, n = None
    assert rel(TR.DST(x[0]), O.dst(x[0])) < 1e-14
    assert rel(TR.IDCT(a, n=40), O.idct(a[:, :40])) < 1e-14   # n < len truncates the series
    with pytest.raises(NotImplementedError):
        TR.IDST(a, n=40)


@pytest.mark.gpu
@pytest.mark.parametrize("sym", [False, True])
def test_operators_dropin(mods, o, sym):
    MO = mods["Matrix_Operators"]
    mask = O.eq_sym_mask(K, NR) if sym else np.ones(3 * N)
    X = GOLD["X_big"] * mask
    psi = X[:N]
    D, R = MO.cheb_radial(C1["N_r"], C1["d"])
    omask = mask[N:2 * N]   # cosine rows kept
    smask = mask[:N]        # sine rows kept
    assert rel(MO.J_theta_RT(psi, NR, K, sym), O.j_theta(psi, K, NR) * omask) < 1e-14
    assert rel(MO.DT0_theta(psi, o.DT0, K, NR, sym), O.dt0_theta(psi, o.DT0, K, NR) * omask) < 1e-14
    assert rel(MO.A2_SINE(psi, D, R, K, NR, sym), O.a2_sine(psi, o) * smask) < 1e-14
    assert rel(MO.A2_SINE_R2(psi, K, NR, D, R, sym), O.a2_sine_r2(psi, o) * smask) < 1e-14
    tag = "sym" if sym else "full"
    assert rel(MO.NLIN_FX(GOLD["X_big"], D, R, K, NR, sym), GOLD["nlin_fx_big_" + tag]) < 1e-13
    assert rel(MO.NLIN_DFX(GOLD["dv"], GOLD["X_big"], D, R, K, NR, sym), GOLD["nlin_dfx_" + tag]) < 1e-13
    if not sym:
        JPSI = MO.J_theta_RT(psi, NR, K, sym)
        OMEGA = MO.A2_SINE_R2(psi, K, NR, D, R, sym)
        slabs = MO.Derivatives(X, JPSI, OMEGA, D[1:-1, 1:-1], K, NR, sym)
        names = ["JT", "kDpsi", "komega", "DT", "DC", "omega", "Dpsi", "kT", "kC"]
        for name, s in zip(names, slabs):
            assert rel(s, GOLD["slab_" + name]) < 1e-14, name
    # back substitutions with the reference's operator tuple
    Main = mods["Main"]
    _, _, _, _, _, LA4, LT, LS, args_A4, _ = Main.Build_Matrix_Operators(
        K, C1["N_r"], C1["d"], C1["dt"], C1["Pr"], C1["Tau"])
    g = np.random.default_rng(4).standard_normal(N)
    assert rel(MO.NAB2_BSub_TSTEP_V2(g, LT, K, NR, C1["dt"], sym),
               O.nab2_bsub(g, o.Linv_T, K, NR, C1["dt"], sym)) < 1e-14
    assert rel(MO.NAB2_BSub_TSTEP_V2(g, LS, K, NR, C1["Tau"] * C1["dt"], sym),
               O.nab2_bsub(g, o.Linv_S, K, NR, C1["Tau"] * C1["dt"], sym)) < 1e-14
    pdt = C1["Pr"] * C1["dt"]
    assert rel(MO.A4_BSub_TSTEP_V2(g, LA4, *args_A4, pdt, sym),
               O.a4_bsub(g, o.Linv_A4, o.D2_A4, o.IR4_A4, o.IR2_A4, K, NR, pdt, sym)) < 1e-14


@pytest.mark.gpu
def test_step_composed_like_reference(mods):
    """Step_Python (Main.py:255-283) written with the drop-in functions and the
    operator tuple, exactly as the reference composes it, against the oracle."""
    Main, MO = mods["Main"], mods["Matrix_Operators"]
    dt, Pr, Tau, Ra, Ra_s = C1["dt"], C1["Pr"], C1["Tau"], C1["Ra"], C1["Ra_s"]
    D, R, Rsq, DT0, gr_k, LA4, LT, LS, args_A4, args_FX = Main.Build_Matrix_Operators(
        K, C1["N_r"], C1["d"], dt, Pr, Tau)
    X = GOLD["X_small"]
    psi, T, S = X[:N], X[N:2 * N], X[2 * N:]
    NX = -1 * dt * MO.NLIN_FX(X, *args_FX, False)
    psi_T0 = MO.DT0_theta(psi, DT0, K, NR, False)
    Om = MO.A2_SINE(psi, D, R, K, NR, False)
    NX[:N] += Om + dt * Pr * gr_k.dot(Ra * T - Ra_s * S)
    psi_new = MO.A4_BSub_TSTEP_V2(NX[:N], LA4, *args_A4, Pr * dt, False)
    NX[N:2 * N] += Rsq.dot(T) - dt * psi_T0
    T_new = MO.NAB2_BSub_TSTEP_V2(NX[N:2 * N], LT, K, NR, dt, False)
    NX[2 * N:] += Rsq.dot(S) - dt * psi_T0
    S_new = MO.NAB2_BSub_TSTEP_V2(NX[2 * N:], LS, K, NR, Tau * dt, False)
    o = O.build_operators(K, C1["N_r"], C1["d"], dt, Pr, Tau)
    assert rel(np.hstack((psi_new, T_new, S_new)), O.step(X, o, Ra, Ra_s)) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("sym", [False, True])
def test_time_step_driver_vs_golden(mods, tmp_path, sym):
    """_Time_Step through the drop-in against the reference driver's golden
    20-step run: final state and the per-step diagnostics it checkpoints."""
    Main = mods["Main"]
    tag = "sym" if sym else "full"
    fname = str(tmp_path / "run.h5")
    Xn = Main._Time_Step(GOLD["X_small"].copy(), C1["Ra"], C1["Ra_s"], C1["Tau"], C1["Pr"],
                         C1["d"], K, C1["N_r"], sym, save_filename=fname, start_time=0,
                         Total_time=20 * C1["dt"], dt=C1["dt"])
    assert rel(Xn, GOLD["traj20_X_" + tag]) < 1e-13
    try:
        import h5py
    except ImportError:
        h5py = None
    if h5py is None:   # the package's own HDF5 writer made the file
        from spectraldoublediffusiveconvection_amd.dropin import _hdf5
        f = _hdf5.read(fname)
        sd = {k: f["Scalar_Data"][k] for k in ("Norm", "KE", "Nu_T", "Nu_S")}
        assert f["Checkpoints"]["X_DATA"].shape[1] == 3 * N
    else:
        with h5py.File(fname, "r") as h:
            sd = {k: np.asarray(h["Scalar_Data"][k]) for k in ("Norm", "KE", "Nu_T", "Nu_S")}
    for k, v in sd.items():
        n = len(v)
        assert n >= 19
        gold = GOLD["traj20_%s_%s" % (k, tag)][:n]
        assert np.max(np.abs(v / gold - 1)) < 1e-12, k


@pytest.mark.gpu
def test_diagnostics_dropin(mods, o):
    Main = mods["Main"]
    X = GOLD["traj20_X_full"]
    ke = Main.Kinetic_Energy(X, o.R, o.D, K, NR, False)
    assert abs(ke / O.kinetic_energy(X, o.R, o.D, K, NR) - 1) < 1e-12
    nu = Main.Nusselt(X[N:2 * N], C1["d"], o.R, o.D, K, NR, check=False)
    assert abs(nu / O.nusselt(X[N:2 * N], C1["d"], o.R, o.D, K, NR, check=False) - 1) < 1e-12
    # the check: the reference raises when Nu(R_1) and Nu(R_2) disagree by > 0.1%
    try:
        O.nusselt(X[N:2 * N], C1["d"], o.R, o.D, K, NR, check=True)
        ref_raises = False
    except ValueError:
        ref_raises = True
    if ref_raises:
        with pytest.raises(ValueError):
            Main.Nusselt(X[N:2 * N], C1["d"], o.R, o.D, K, NR, check=True)
    else:
        Main.Nusselt(X[N:2 * N], C1["d"], o.R, o.D, K, NR, check=True)


@pytest.mark.gpu
def test_newton_dropin_steady_convection(mods):
    """The reference's wide-gap Newton case (Run_Tests.py:140-170, 201-235):
    time-step to steady convection (on the GPU), then _Newton from there.  The
    converged state is a fixed point of the oracle's step at dt = 1 and agrees
    with the oracle's own Newton solve from the same start."""
    import torch
    from spectraldoublediffusiveconvection_amd import device as dev
    Main = mods["Main"]
    d, Kw, Nr, dt = 2.0, 32, 16, 0.075
    Ra, Ra_s, Tau, Pr = 6780.0, 0.0, 1.0, 10.0
    n3 = 3 * Kw * (Nr - 1)
    X0 = np.random.default_rng(6).random(n3)
    X0 = 1e-3 * X0 / np.linalg.norm(X0)
    plan = dev.Plan.build(Kw, Nr, d, dt, Pr, Tau, symmetric=True)
    X_ts = plan.step(torch.as_tensor(X0, device="cuda"), Ra, Ra_s, nsteps=26667)
    torch.cuda.synchronize()
    X_ts = X_ts.cpu().numpy()[0]
    assert np.linalg.norm(X_ts) > 1e-2          # convection, not the conductive state
    out = Main._Newton(X_ts, Ra, Ra_s, Tau, Pr, d, Kw, Nr, True, tol_newton=1e-8)
    assert out[-1] is True
    Xs = out[0]
    o = O.build_operators(Kw, Nr, d, 1.0, Pr, Tau)
    assert rel(O.step(Xs, o, Ra, Ra_s, symmetric=True), Xs) < 1e-9
    ref = O.newton(X_ts, Ra, Ra_s, Tau, Pr, d, Kw, Nr, True, tol_newton=1e-8)
    assert ref[-1] is True
    assert rel(Xs, ref[0]) < 1e-8
    assert abs(out[2] / ref[2] - 1) < 1e-8 and abs(out[3] / ref[3] - 1) < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("K_o,K_n", [(16, 32), (32, 16), (48, 64), (64, 48)])
def test_interp_thetas_vs_oracle_transforms(mods, K_o, K_n):
    """INTERP_THETAS (device transforms) against the oracle's transforms
    composed as Matrix_Operators.py:944-1011 composes them: fields laid out
    on max(K_o, K_n) columns, IDST/IDCT onto that grid, DST/DCT truncated to
    K_n modes.  Plus: a band-limited state that is carried up and back down
    comes back, every row but psi's first: as in the reference, psi's rows
    enter the sine transforms without the shift into the sinusoid index, so
    row 0 sits in the sin(0 theta) slot and is dropped."""
    MO = mods["Matrix_Operators"]
    nr = 7
    X_o = np.random.default_rng(K_o + K_n).standard_normal(3 * nr * K_o)
    got = MO.INTERP_THETAS(K_n, K_o, X_o)
    width = max(K_o, K_n)
    F = [np.zeros((nr, width)) for _ in range(3)]
    for f in range(3):
        F[f][:, :K_o] = X_o[f * K_o * nr:(f + 1) * K_o * nr].reshape(K_o, nr).T
    hats = (O.dst(O.idst(F[0]), n=K_n), O.dct(O.idct(F[1]), n=K_n), O.dct(O.idct(F[2]), n=K_n))
    want = np.concatenate([h.T.ravel() for h in hats])
    assert got.shape == want.shape
    assert rel(got, want) < 1e-14
    assert MO.INTERP_THETAS(K_o, K_o, X_o) is X_o
    if K_n > K_o:
        back = MO.INTERP_THETAS(K_o, K_n, got)
        keep = np.ones(3 * nr * K_o, dtype=bool)
        keep[:nr] = False
        assert rel(back[keep], X_o[keep]) < 1e-13
        assert not back[:nr].any()
