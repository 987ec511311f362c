"""GPU parity of libsddc against the CPU oracle and the reference's golden
vectors (float64 plain method), plus the mixed-precision bound and the
size-independent properties at the full workload size.

Tolerances (relative 2-norm unless stated):
  * transforms, theta operators, slabs, solves: 1e-14
  * T and S solves in the eigenbasis of r^2 - h r^2 Lap (nab2_solve 2):
    5e-14 (the eigenvectors' condition number is about 2; measured up to
    1.1e-14 at N_r = 65)
  * nonlinear terms F(X,X) and F(X,dv)+F(dv,X):   1e-13
  * one step:                                     1e-12 (the per-step bound)
  * kinetic energy / Nusselt over a trajectory:   1e-9
  * mixed precision: per-step state error <= 1e-5, kinetic energy <= 1e-6
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import sddc_oracle as O  # noqa: E402
from spectraldoublediffusiveconvection_amd import _lib  # noqa: E402
from spectraldoublediffusiveconvection_amd import device as dev  # noqa: E402
from spectraldoublediffusiveconvection_amd import operators as ops  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = dict(np.load(os.path.join(HERE, "golden", "golden_c1.npz")))
C1 = dict(N_r=20, N_fm=64, d=0.35, Pr=1.0, Tau=1.0 / 15.0, Ra=9000.0, Ra_s=500.0, dt=5e-3)
# Long runs use Ra_T = 3000: at Ra_T = 9000 the conductive state grows like
# exp(~20 t) and every resolution overflows within ~500 steps (the reference
# algorithm included); Ra_T = 3000 stays bounded for 10^4 steps.
RA_LONG = 3000.0
TOL_NAB2 = 5e-14


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


def H(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def c1():
    o = O.build_operators(C1["N_fm"], C1["N_r"], C1["d"], C1["dt"], C1["Pr"], C1["Tau"])
    p = dev.Plan.build(C1["N_fm"], C1["N_r"], C1["d"], C1["dt"], C1["Pr"], C1["Tau"], max_batch=4)
    return o, p


def _transform(kind, a, rows, len_in, n_out):
    out = torch.empty((rows, n_out), dtype=torch.float64, device="cuda")
    src = T(a)
    _lib.check(_lib.lib().sddc_transform(kind, ctypes.c_void_p(src.data_ptr()),
                                         ctypes.c_void_p(out.data_ptr()), rows, len_in, n_out,
                                         None))
    return H(out)


@pytest.mark.parametrize("K", [64, 10, 16, 48, 256, 512])
def test_transforms(K):
    M = 3 * K // 2
    rng = np.random.default_rng(K)
    a = rng.standard_normal((7, K)) * np.logspace(0, 3, 7)[:, None]   # rows of unequal scale
    x = rng.standard_normal((7, M)) * np.logspace(3, 0, 7)[:, None]
    assert rel(_transform(_lib.SDDC_IDCT, a, 7, K, M), O.idct(a, M)) < 1e-14
    assert rel(_transform(_lib.SDDC_IDST, a, 7, K, M), O.idst(a, M)) < 1e-14
    assert rel(_transform(_lib.SDDC_DCT, x, 7, M, K), O.dct(x)[:, :K]) < 1e-14
    assert rel(_transform(_lib.SDDC_DST, x, 7, M, K), O.dst(x)[:, :K]) < 1e-14
    # n = None (same length) and the reference KAT sizes
    assert rel(_transform(_lib.SDDC_IDCT, a, 7, K, K), O.idct(a)) < 1e-14
    assert rel(_transform(_lib.SDDC_DST, a, 7, K, K), O.dst(a)) < 1e-14


def test_theta_ops_and_derivatives(c1):
    o, p = c1
    K, nr = C1["N_fm"], C1["N_r"] - 1
    X = GOLD["X_big"]
    t = H(p.theta_ops(T(X)))[0]
    psi = X[:K * nr]
    assert rel(t[0], O.j_theta(psi, K, nr)) < 1e-14
    assert rel(t[1], O.dt0_theta(psi, o.DT0, K, nr)) < 1e-14
    assert rel(t[2], O.a2_sine(psi, o)) < 1e-14
    assert rel(t[3], O.a2_sine_r2(psi, o)) < 1e-14
    s = H(p.derivatives(T(X)))[0]
    for i, n in enumerate(["JT", "kDpsi", "komega", "DT", "DC", "omega", "Dpsi", "kT", "kC"]):
        assert rel(s[i], GOLD["slab_" + n]) < 1e-14, n


def test_nonlinear_terms_vs_golden(c1):
    _, p = c1
    assert rel(H(p.nlin_fx(T(GOLD["X_big"])))[0], GOLD["nlin_fx_big_full"]) < 1e-13
    assert rel(H(p.nlin_fx(T(GOLD["X_small"])))[0], GOLD["nlin_fx_small_full"]) < 1e-13
    assert rel(H(p.nlin_dfx(T(GOLD["dv"]), T(GOLD["X_big"])))[0], GOLD["nlin_dfx_full"]) < 1e-13


def test_back_substitutions(c1):
    o, p = c1
    K, nr = C1["N_fm"], C1["N_r"] - 1
    g = np.random.default_rng(1).standard_normal((2, K * nr))
    a4 = H(p.bsub(T(g), 0))
    t = H(p.bsub(T(g), 1))
    s = H(p.bsub(T(g), 2))
    for b in range(2):
        assert rel(a4[b], O.a4_bsub(g[b], o.Linv_A4, o.D2_A4, o.IR4_A4, o.IR2_A4, K, nr,
                                    C1["Pr"] * C1["dt"])) < 1e-14
        assert rel(t[b], O.nab2_bsub(g[b], o.Linv_T, K, nr, C1["dt"])) < 1e-14
        assert rel(s[b], O.nab2_bsub(g[b], o.Linv_S, K, nr, C1["Tau"] * C1["dt"])) < 1e-14


@pytest.mark.parametrize("K,Nr", [(64, 20), (512, 40), (10, 9)])
@pytest.mark.parametrize("max_batch", [1, 16, 128])
@pytest.mark.parametrize("sym", [False, True])
def test_solve_segmentations(K, Nr, max_batch, sym):
    """The chain solves of every segmentation a plan can pick (rows per
    segment ~sqrt(n), ~2 sqrt(n), one segment; n = K/2 rows per chain) against
    the oracle's sequential back substitution, for a ragged batch of 5 (padded
    workgroups), then three steps (the solve's psi suffix sums feed the next
    step)."""
    d, dt, Pr, Tau = 0.35, 5e-3, 1.0, 1 / 15.0
    o = O.build_operators(K, Nr, d, dt, Pr, Tau)
    p = dev.Plan(ops.build(K, Nr, d, dt, Pr, Tau), symmetric=sym, max_batch=max_batch)
    nr = Nr - 1
    rng = np.random.default_rng(K + max_batch)
    B = min(5, max_batch)
    g = rng.standard_normal((B, K * nr)) * np.logspace(0, -3, K).repeat(nr)[None, :]
    dts = (Pr * dt, dt, Tau * dt)
    for field in range(3):
        f = H(p.bsub(T(g), field))
        for b in range(B):
            if field == 0:
                ref = O.a4_bsub(g[b], o.Linv_A4, o.D2_A4, o.IR4_A4, o.IR2_A4, K, nr, dts[0], sym)
            else:
                Linv = o.Linv_T if field == 1 else o.Linv_S
                ref = O.nab2_bsub(g[b], Linv, K, nr, dts[field], sym)
            assert rel(f[b], ref) < 1e-14, (field, b)
    X = 1e-3 * rng.standard_normal((B, 3 * K * nr))
    Y = H(p.step(T(X), RA_LONG, 500.0, nsteps=3))
    for b in range(B):
        ref = X[b]
        for _ in range(3):
            ref = O.step(ref, o, RA_LONG, 500.0, symmetric=sym)
        assert rel(Y[b], ref) < 1e-12, b


@pytest.mark.parametrize("K,Nr", [(64, 20), (512, 40), (10, 9), (32, 65)])
@pytest.mark.parametrize("segment", [0, 8])
@pytest.mark.parametrize("sym", [False, True])
@pytest.mark.parametrize("width", [16, 3, 5, 6, 9, 10])
def test_solve_mfma(K, Nr, segment, sym, width):
    """The 16-instance MFMA chain solve (solve_instances = 16) and the
    mixed-width quad solves (3, 5, 6: 2 or 3 instances per A4 workgroup, 4
    per NAB2 workgroup; 9, 10: one per A4 workgroup with deeper prefetch)
    against the oracle's back substitution, for a ragged batch of
    21 (partial workgroups), unsegmented and in segments of 8 rows, then
    three steps against the oracle."""
    d, dt, Pr, Tau = 0.35, 5e-3, 1.0, 1 / 15.0
    o = O.build_operators(K, Nr, d, dt, Pr, Tau)
    B = 21
    p = dev.Plan(ops.build(K, Nr, d, dt, Pr, Tau), symmetric=sym, max_batch=B,
                 tuning={"chain_segment": segment, "solve_instances": width})
    assert p.tuning()["solve_instances"] == width
    nr = Nr - 1
    rng = np.random.default_rng(K + segment)
    g = rng.standard_normal((B, K * nr)) * np.logspace(0, -3, K).repeat(nr)[None, :]
    dts = (Pr * dt, dt, Tau * dt)
    for field in range(3):
        f = H(p.bsub(T(g), field))
        for b in range(B):
            if field == 0:
                ref = O.a4_bsub(g[b], o.Linv_A4, o.D2_A4, o.IR4_A4, o.IR2_A4, K, nr, dts[0], sym)
            else:
                Linv = o.Linv_T if field == 1 else o.Linv_S
                ref = O.nab2_bsub(g[b], Linv, K, nr, dts[field], sym)
            assert rel(f[b], ref) < 1e-14, (field, b)
    X = 1e-3 * rng.standard_normal((B, 3 * K * nr))
    Y = H(p.step(T(X), RA_LONG, 500.0, nsteps=3))
    for b in (0, 7, 15, 16, 20):
        ref = X[b]
        for _ in range(3):
            ref = O.step(ref, o, RA_LONG, 500.0, symmetric=sym)
        assert rel(Y[b], ref) < 1e-12, b


@pytest.mark.parametrize("K,Nr", [(64, 20), (512, 40), (10, 9), (32, 65)])
@pytest.mark.parametrize("sym", [False, True])
def test_nab2_eigenbasis_solve(K, Nr, sym):
    """The T and S solves in the eigenbasis (nab2_solve 2) against the
    per-mode block chains (the default) and the oracle's NAB2_BSub_TSTEP_V2, for a ragged
    batch of 5 over chunked chains (K = 512: four chunks of 64 rows per
    chain; N_r = 65: the largest radial size), then three steps of both
    plans against each other and the oracle."""
    d, dt, Pr, Tau = 0.35, 5e-3, 1.0, 1 / 15.0
    o = O.build_operators(K, Nr, d, dt, Pr, Tau)
    op = ops.build(K, Nr, d, dt, Pr, Tau)
    B = 5
    pm = dev.Plan(op, symmetric=sym, max_batch=B, tuning={"nab2_solve": 2})
    pb = dev.Plan(op, symmetric=sym, max_batch=B)
    assert pm.tuning()["nab2_solve"] == 2 and pb.tuning()["nab2_solve"] == 1
    nr = Nr - 1
    rng = np.random.default_rng(K + Nr)
    g = rng.standard_normal((B, K * nr)) * np.logspace(0, -3, K).repeat(nr)[None, :]
    for field, Linv, h in ((1, o.Linv_T, dt), (2, o.Linv_S, Tau * dt)):
        fm, fb = H(pm.bsub(T(g), field)), H(pb.bsub(T(g), field))
        for b in range(B):
            ref = O.nab2_bsub(g[b], Linv, K, nr, h, sym)
            assert rel(fm[b], ref) < TOL_NAB2, (field, b)
            assert rel(fb[b], ref) < 1e-14, (field, b)
    X = 1e-3 * rng.standard_normal((B, 3 * K * nr))
    Ym = H(pm.step(T(X), RA_LONG, 500.0, nsteps=3))
    Yb = H(pb.step(T(X), RA_LONG, 500.0, nsteps=3))
    for b in range(B):
        ref = X[b]
        for _ in range(3):
            ref = O.step(ref, o, RA_LONG, 500.0, symmetric=sym)
        assert rel(Ym[b], ref) < 1e-12, b
        assert rel(Ym[b], Yb[b]) < 1e-13, b
    pm.close()
    pb.close()


@pytest.mark.skip(reason="capturing a graph of an eigenbasis plan, with that solve forked onto "
                         "a side stream per sub-batch, ended the process with a segmentation fault "
                         "inside sddc_graph_create on MI355X; the cause was not found, the fork "
                         "was removed and graph capture now refuses such plans")
def test_nab2_eigenbasis_beside_the_a4_chains_in_graph_parts():
    """The eigenbasis solve runs on a side stream beside the A4 chains: a
    graph of two sub-batches (each with its own side stream) replays to the
    bits of the eager steps, twice, and both agree with the block form to
    rounding."""
    K, Nr, d, dt, Pr, Tau = 64, 20, 0.35, 5e-3, 1.0, 1 / 15.0
    op = ops.build(K, Nr, d, dt, Pr, Tau)
    B = 32
    pm = dev.Plan(op, max_batch=B, tuning={"nab2_solve": 2, "sub_batches": 2})
    pb = dev.Plan(op, max_batch=B)
    X = 1e-3 * np.random.default_rng(3).standard_normal((B, 3 * K * (Nr - 1)))
    Ra = np.linspace(2500.0, 3000.0, B)
    eager = T(X)
    xg = T(X).clone()
    gr = pm.graph(xg, Ra, 500.0, nsteps=3)
    for rep in range(2):
        gr.launch()
        eager = pm.step(eager, Ra, 500.0, nsteps=3)
        assert np.array_equal(H(xg), H(eager)), rep
    gr.close()
    ref = H(pb.step(T(X), Ra, 500.0, nsteps=6))
    assert rel(H(eager), ref) < 1e-12
    pm.close()
    pb.close()


def test_graph_capture_refuses_eigenbasis_plans():
    """Graph capture checks its plans before it captures anything: a plan
    with nab2_solve 2 is refused with NotImplementedError."""
    op = ops.build(64, 20, 0.35, 5e-3, 1.0, 1 / 15.0)
    pm = dev.Plan(op, max_batch=4, tuning={"nab2_solve": 2})
    X = T(1e-3 * np.random.default_rng(4).standard_normal((4, 3 * 64 * 19)))
    with pytest.raises(NotImplementedError, match="nab2_solve 2"):
        pm.graph(X, 2500.0, 500.0, nsteps=2)
    pm.close()


def test_nab2_eigenbasis_needs_the_modes():
    """An operator set without modes_T/modes_S (the drop-ins' caller-supplied
    matrices) solves with the blocks; asking for the eigenbasis is refused."""
    o = ops.build(16, 12, 0.35, 5e-3, 1.0, 1.0)
    bare = ops.from_matrices(o.D, o.R, 16, dt=5e-3, Linv_A4=o.Linv_A4, Linv_T=o.Linv_T,
                             Linv_S=o.Linv_S)
    assert bare.modes_T is None
    assert dev.Plan(bare).tuning()["nab2_solve"] == 1
    with pytest.raises(ValueError, match="nab2_solve 2"):
        dev.Plan(bare, tuning={"nab2_solve": 2})
    with pytest.raises(ValueError, match="nab2_solve must be"):
        dev.Plan(o, tuning={"nab2_solve": 3})


@pytest.mark.parametrize("group", [(1, 2, 4, 10), (5, 7, 8, 9)])
def test_solve_width_groups_give_the_same_bits(group):
    """Widths that differ only in how many instances share a workgroup or in
    how many rows of operator data are in flight do the same arithmetic per
    instance: the steps of a ragged batch agree bit for bit within each
    group (unsegmented and in segments of 8 rows)."""
    K, Nr, d, dt, Pr, Tau = 64, 20, 0.35, 5e-3, 1.0, 1 / 15.0
    op = ops.build(K, Nr, d, dt, Pr, Tau)
    rng = np.random.default_rng(sum(group))
    B = 7
    X = 1e-3 * rng.standard_normal((B, 3 * K * (Nr - 1)))
    Ra = np.linspace(2500.0, 9000.0, B)
    for seg in (0, 8):
        outs = []
        for w in group:
            p = dev.Plan(op, max_batch=B, tuning={"chain_segment": seg, "solve_instances": w})
            outs.append(H(p.step(T(X), Ra, 500.0, nsteps=3)))
            p.close()
        for w, o in zip(group[1:], outs[1:]):
            assert np.array_equal(o, outs[0]), (seg, w)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spectral_instances_give_the_same_bits(precision):
    """Instances per spectral workgroup (1-4, and the automatic choice) only
    change which workgroup computes an instance: a ragged batch of 7 steps
    to the same bits with each."""
    K, Nr, d, dt, Pr, Tau = 128, 20, 0.35, 5e-3, 1.0, 1 / 15.0
    op = ops.build(K, Nr, d, dt, Pr, Tau)
    B = 7
    X = 1e-3 * np.random.default_rng(5).standard_normal((B, 3 * K * (Nr - 1)))
    outs = []
    for sp in (0, 1, 2, 3, 4):
        p = dev.Plan(op, precision=precision, max_batch=B, tuning={"spectral_instances": sp})
        assert p.tuning()["spectral_instances"] == sp
        outs.append(H(p.step(T(X), np.linspace(2500.0, 3000.0, B), 500.0, nsteps=3)))
        p.close()
    for sp, o in zip((1, 2, 3, 4), outs[1:]):
        assert np.array_equal(o, outs[0]), sp


@pytest.mark.parametrize("width", [3, 5, 6, 7])
@pytest.mark.parametrize("seg", [8, 256])
def test_mixed_width_launch_covers_every_instance(width, seg):
    """The mixed-width solves (widths 3, 5-10) launch a one-dimensional grid
    of their live (field, chain, segment, instance group) items, cut into
    eight runs by XCD and padded to a multiple of eight.  For batches whose
    item counts are not multiples of eight, one or four segments, every
    instance of every field solves to the same bits as that instance alone,
    and instance 0 to the oracle within 1e-14."""
    K, Nr, d, dt, Pr, Tau = 64, 20, 0.35, 5e-3, 1.0, 1 / 15.0
    o = O.build_operators(K, Nr, d, dt, Pr, Tau)
    op = ops.build(K, Nr, d, dt, Pr, Tau)
    nr = Nr - 1
    rng = np.random.default_rng(100 * width + seg)
    tuning = {"chain_segment": seg, "solve_instances": width}
    alone = dev.Plan(op, max_batch=1, tuning=tuning)
    for B in (5, 13, 37):
        p = dev.Plan(op, max_batch=B, tuning=tuning)
        g = rng.standard_normal((B, K * nr))
        for field in range(3):
            f = H(p.bsub(T(g), field))
            for b in range(B):
                assert np.array_equal(f[b], H(alone.bsub(T(g[b:b + 1]), field))[0]), (B, field, b)
            if field == 0:
                ref = O.a4_bsub(g[0], o.Linv_A4, o.D2_A4, o.IR4_A4, o.IR2_A4, K, nr, Pr * dt, False)
            else:
                ref = O.nab2_bsub(g[0], o.Linv_T if field == 1 else o.Linv_S, K, nr,
                                  dt if field == 1 else Tau * dt, False)
            assert rel(f[0], ref) < 1e-14, (B, field)
        p.close()
    alone.close()


@pytest.mark.parametrize("width", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 16])
def test_solve_widths_padded_instances(width):
    """Regression test: batches that do not fill the last workgroup of a
    width (B = 4 and 5; with three instances per A4 workgroup, instances 4
    and 5 of B = 4 are padding) and sub-batched graphs whose parts do not
    divide by the width, in segments of 8 rows (so the joins read the
    segment exit states).  Padded instances write the exit states of the
    last real one; before, they wrote past the exit-state workspace or into
    the next sub-batch's."""
    K, Nr, d, dt, Pr, Tau = 64, 20, 0.35, 5e-3, 1.0, 1 / 15.0
    o = O.build_operators(K, Nr, d, dt, Pr, Tau)
    op = ops.build(K, Nr, d, dt, Pr, Tau)
    nr = Nr - 1
    rng = np.random.default_rng(width)
    dts = (Pr * dt, dt, Tau * dt)
    for B in (4, 5):
        p = dev.Plan(op, max_batch=B, tuning={"chain_segment": 8, "solve_instances": width})
        g = rng.standard_normal((B, K * nr))
        for field in range(3):
            f = H(p.bsub(T(g), field))
            for b in range(B):
                if field == 0:
                    ref = O.a4_bsub(g[b], o.Linv_A4, o.D2_A4, o.IR4_A4, o.IR2_A4, K, nr, dts[0],
                                    False)
                else:
                    ref = O.nab2_bsub(g[b], o.Linv_T if field == 1 else o.Linv_S, K, nr,
                                      dts[field], False)
                assert rel(f[b], ref) < 1e-14, (B, field, b)
        p.close()
    B = 64
    p = dev.Plan(op, max_batch=B, tuning={"chain_segment": 8, "solve_instances": width,
                                          "sub_batches": 2})
    X = 1e-3 * rng.standard_normal((B, 3 * K * nr))
    Ra = np.linspace(2500.0, 9000.0, B)
    eager = T(X)
    xg = T(X).clone()
    gr = p.graph(xg, Ra, 500.0, nsteps=3)   # two parts of 32 on forked streams
    for rep in range(2):
        gr.launch()
        eager = p.step(eager, Ra, 500.0, nsteps=3)
        assert np.array_equal(H(xg), H(eager)), rep
    gr.close()
    for b in (0, 31, 32, 63):
        ref = X[b]
        for _ in range(3):
            ref = O.step(ref, o, Ra[b], 500.0)
        assert rel(H(p.step(T(X[b:b + 1]), Ra[b], 500.0, nsteps=3))[0], ref) < 1e-12, b
    p.close()


@pytest.mark.parametrize("K,Nr,segment", [(512, 40, 0), (64, 20, 8), (64, 40, 8), (32, 65, 8)])
def test_solve_width_groups_same_bits(K, Nr, segment):
    """Solve widths 1-4 give the same bits, as do 5 and 6: within a group the
    per-row arithmetic and summation order are the same, and the widths differ
    only in how many instances share a workgroup and how many rows of operator
    data are in flight (three for width 5, two for the others).  At N_r = 40
    and 65 width 4's segment joins take two instances per workgroup instead
    of four (four need one wavefront per SIMD or spill there)."""
    op = ops.build(K, Nr, 0.35, 5e-3, 1.0, 1 / 15.0)
    nr = Nr - 1
    rng = np.random.default_rng(K)
    B = 7
    g = rng.standard_normal((B, K * nr))
    X = 1e-3 * rng.standard_normal((B, 3 * K * nr))
    Ra = np.linspace(2500.0, 3000.0, B)
    outs = {}
    for width in (1, 2, 3, 4, 5, 6):
        p = dev.Plan(op, max_batch=B, tuning={"chain_segment": segment, "solve_instances": width})
        outs[width] = [H(p.bsub(T(g), f)) for f in range(3)] + [H(p.step(T(X), Ra, 500.0, nsteps=2))]
        p.close()
    for group in ((1, 2, 3, 4), (5, 6)):
        for w in group[1:]:
            for i, (a, b) in enumerate(zip(outs[group[0]], outs[w])):
                assert np.array_equal(a, b), (group[0], w, i)


def test_single_step_and_linearised(c1):
    o, p = c1
    for name in ("X_big", "X_small"):
        X = GOLD[name]
        assert rel(H(p.step(T(X), C1["Ra"], C1["Ra_s"]))[0], O.step(X, o, C1["Ra"], C1["Ra_s"])) < 1e-12
        assert rel(H(p.step(T(X), C1["Ra"], C1["Ra_s"], linear=True))[0],
                   O.step(X, o, C1["Ra"], C1["Ra_s"], linear=True)) < 1e-12
    a = H(p.step_linearized(T(GOLD["dv"]), T(GOLD["X_big"]), C1["Ra"], C1["Ra_s"]))[0]
    assert rel(a, O.step_linearized(GOLD["dv"], GOLD["X_big"], o, C1["Ra"], C1["Ra_s"])) < 1e-12


@pytest.mark.parametrize("sym", [False, True])
def test_golden_trajectory(sym):
    """20 steps of the reference driver (_Time_Step) incl. the symmetry mask."""
    p = dev.Plan.build(C1["N_fm"], C1["N_r"], C1["d"], C1["dt"], C1["Pr"], C1["Tau"], symmetric=sym)
    tag = "sym" if sym else "full"
    X = H(p.step(T(GOLD["X_small"]), C1["Ra"], C1["Ra_s"], nsteps=20))[0]
    assert rel(X, GOLD["traj20_X_" + tag]) < 1e-13
    d = H(p.diagnostics(T(X)))[0]
    for j, key in enumerate(("Norm", "KE", "Nu_T", "Nu_S")):
        assert abs(d[j] / GOLD["traj20_%s_%s" % (key, tag)][-1] - 1) < 1e-12, key
    if sym:
        assert rel(H(p.nlin_fx(T(GOLD["X_big"])))[0], GOLD["nlin_fx_big_sym"]) < 1e-13
        assert rel(H(p.nlin_dfx(T(GOLD["dv"]), T(GOLD["X_big"])))[0], GOLD["nlin_dfx_sym"]) < 1e-13


def test_large_state_trajectory_vs_golden(c1):
    _, p = c1
    X = H(p.step(T(GOLD["X_big"]), C1["Ra"], C1["Ra_s"], nsteps=10))[0]
    assert rel(X, GOLD["traj10big_X_full"]) < 1e-12
    X = H(p.step(T(GOLD["X_small"]), C1["Ra"], C1["Ra_s"], nsteps=10, linear=True))[0]
    assert rel(X, GOLD["traj10lin_X_full"]) < 1e-13


def test_thousand_steps_per_step_parity(c1):
    """The CPU-runnable workload: 1000 steps from a seeded N(0,1e-3) state,
    every step compared with the oracle's step of the same input, and the
    kinetic energy / Nusselt numbers of the final states."""
    o, p = c1
    X0 = np.random.default_rng(2024).standard_normal(3 * C1["N_fm"] * (C1["N_r"] - 1))
    X0 = 1e-3 * X0 / np.linalg.norm(X0)
    xg = T(X0).view(1, -1)
    xc = X0.copy()
    worst = 0.0
    for it in range(1000):
        x_in = H(xg)[0] if it % 10 == 0 else None
        xg = p.step(xg, RA_LONG, C1["Ra_s"])
        if x_in is not None:   # per-step error: the oracle's step of the GPU's own input
            worst = max(worst, rel(H(xg)[0], O.step(x_in, o, RA_LONG, C1["Ra_s"])))
        xc = O.step(xc, o, RA_LONG, C1["Ra_s"])
    assert worst < 1e-12, worst
    assert rel(H(xg)[0], xc) < 1e-9
    d = H(p.diagnostics(xg))[0]
    K, nr = C1["N_fm"], C1["N_r"] - 1
    N = K * nr
    ke = O.kinetic_energy(xc, o.R, o.D, K, nr)
    assert abs(d[1] / ke - 1) < 1e-9
    assert abs(d[2] / O.nusselt(xc[N:2 * N], C1["d"], o.R, o.D, K, nr, check=False) - 1) < 1e-9
    assert abs(d[3] / O.nusselt(xc[2 * N:], C1["d"], o.R, o.D, K, nr, check=False) - 1) < 1e-9


def test_ten_thousand_steps_vs_oracle(c1):
    """10^4 float64 steps of the CPU-runnable workload on the GPU and in the
    oracle from the same seeded state: kinetic energy and both Nusselt
    numbers within 1e-9 relative every 1000 steps."""
    o, p = c1
    K, nr = C1["N_fm"], C1["N_r"] - 1
    N = K * nr
    X0 = np.random.default_rng(77).standard_normal(3 * N)
    X0 = 1e-3 * X0 / np.linalg.norm(X0)
    xg = T(X0).view(1, -1)
    xc = X0.copy()
    for blk in range(10):
        xg = p.step(xg, RA_LONG, C1["Ra_s"], nsteps=1000)
        for _ in range(1000):
            xc = O.step(xc, o, RA_LONG, C1["Ra_s"])
        d = H(p.diagnostics(xg))[0]
        ke = O.kinetic_energy(xc, o.R, o.D, K, nr)
        nut = O.nusselt(xc[N:2 * N], C1["d"], o.R, o.D, K, nr, check=False)
        nus = O.nusselt(xc[2 * N:], C1["d"], o.R, o.D, K, nr, check=False)
        assert abs(d[1] / ke - 1) < 1e-9, (blk, d[1], ke)
        assert abs(d[2] / nut - 1) < 1e-9, (blk, d[2], nut)
        assert abs(d[3] / nus - 1) < 1e-9, (blk, d[3], nus)


@pytest.mark.parametrize("K,Nr", [(256, 30), (512, 40), (128, 20), (64, 16), (48, 24), (10, 20)])
def test_step_other_sizes(K, Nr):
    """Every nonlinear kernel variant (in place for K = 128, 256, 512; the
    ping-pong form otherwise) and the step against the oracle in float64, and
    the float32 nonlinear pipeline within its rounding."""
    d, dt, Pr, Tau = 0.35, 5e-3, 1.0, 1 / 15.0
    o = O.build_operators(K, Nr, d, dt, Pr, Tau)
    op = ops.build(K, Nr, d, dt, Pr, Tau)
    p = dev.Plan(op)
    X0 = 1e-3 * np.random.default_rng(K + Nr).standard_normal(3 * K * (Nr - 1))
    assert rel(H(p.step(T(X0), 9000.0, 500.0))[0], O.step(X0, o, 9000.0, 500.0)) < 1e-12
    F = O.nlin_fx(X0, o)
    assert rel(H(p.nlin_fx(T(X0)))[0], F) < 1e-12
    pm = dev.Plan(op, precision="mixed")
    assert rel(H(pm.nlin_fx(T(X0)))[0], F) < 2e-5
    dv = 1e-3 * np.random.default_rng(K).standard_normal(3 * K * (Nr - 1))
    D = O.nlin_dfx(dv, X0, o)
    assert rel(H(p.nlin_dfx(T(dv), T(X0)))[0], D) < 1e-12
    assert rel(H(pm.nlin_dfx(T(dv), T(X0)))[0], D) < 2e-5
    assert rel(H(p.step_linearized(T(dv), T(X0), 9000.0, 500.0))[0],
               O.step_linearized(dv, X0, o, 9000.0, 500.0)) < 1e-12


def test_batched_equals_alone_bitwise(c1):
    """Instances of one size class stepped together give the same bits as alone."""
    _, p = c1
    rng = np.random.default_rng(7)
    X = 1e-3 * rng.standard_normal((3, 3 * C1["N_fm"] * (C1["N_r"] - 1)))
    Ra = np.array([8000.0, 9000.0, 11000.0])
    Ras = np.array([300.0, 500.0, 300.0])
    both = H(p.step(T(X), Ra, Ras, nsteps=5))
    for b in range(3):
        alone = H(p.step(T(X[b]), Ra[b], Ras[b], nsteps=5))[0]
        assert np.array_equal(both[b], alone)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spectral_instance_runs_same_bits(precision):
    """The spectral stage walks runs of up to four instances per workgroup
    (32 mode tiles at N_fm = 512: one instance at B = 18, two at B = 35, four
    at B = 70, each with a partial last run), staging every instance's Ra_T
    and Ra_S and prefetching the next one's state.  Every instance of the
    batched step equals the same instance stepped alone, bit for bit."""
    K, Nr = 512, 20
    p = dev.Plan(ops.build(K, Nr, 0.35, 5e-3, 1.0, 1 / 15.0), precision=precision, max_batch=70)
    rng = np.random.default_rng(21)
    for B in (18, 35, 70):
        X = 1e-3 * rng.standard_normal((B, 3 * K * (Nr - 1)))
        Ra = np.linspace(2500.0, 9000.0, B)
        Ras = np.linspace(100.0, 700.0, B)[::-1].copy()
        both = H(p.step(T(X), Ra, Ras, nsteps=2))
        for b in (0, 1, B // 2, B - 2, B - 1):
            alone = H(p.step(T(X[b]), Ra[b], Ras[b], nsteps=2))[0]
            assert np.array_equal(both[b], alone), (B, b)


def test_graph_replay_matches_eager(c1):
    _, p = c1
    X0 = 1e-3 * np.random.default_rng(9).standard_normal(3 * C1["N_fm"] * (C1["N_r"] - 1))
    xg = T(X0).view(1, -1).clone()
    g = p.graph(xg, C1["Ra"], C1["Ra_s"], nsteps=4)
    eager = T(X0).view(1, -1)
    for rep in range(3):
        g.launch()
        eager = p.step(eager, C1["Ra"], C1["Ra_s"], nsteps=4)
        assert np.array_equal(H(xg), H(eager)), rep
    g.close()


def test_mixed_precision_error_bound():
    """Mixed precision at the full workload size N_r=40, N_fm=512 (float32
    nonlinear pipeline, float64 linear terms and solves) over 10^4 steps:
    kinetic energy within 1e-6 of the float64 run at every 100 steps (the
    target), Nu_T - 1 and Nu_S - 1 within 1e-5 (measured: 4e-8, 6e-7, 1.3e-6)."""
    K, Nr, d, dt, Pr, Tau = 512, 40, 0.35, 5e-3, 1.0, 1 / 15.0
    op = ops.build(K, Nr, d, dt, Pr, Tau)
    p64 = dev.Plan(op)
    p32 = dev.Plan(op, precision="mixed")
    X0 = np.random.default_rng(11).standard_normal(3 * K * (Nr - 1))
    X0 = 1e-3 * X0 / np.linalg.norm(X0)
    x64 = T(X0).view(1, -1)
    x32 = x64.clone()
    worst = np.zeros(3)
    for blk in range(100):
        x64 = p64.step(x64, RA_LONG, 500.0, nsteps=100)
        x32 = p32.step(x32, RA_LONG, 500.0, nsteps=100)
        d64 = H(p64.diagnostics(x64))[0]
        d32 = H(p64.diagnostics(x32))[0]
        assert np.all(np.isfinite(d64))
        worst = np.maximum(worst, np.abs(d32[1:4] / d64[1:4] - 1))
    assert worst[0] < 1e-6, worst
    assert worst[1] < 1e-5 and worst[2] < 1e-5, worst
    big = T(1e-2 * np.random.default_rng(12).standard_normal(3 * K * (Nr - 1))).view(1, -1)
    assert rel(H(p32.step(big, RA_LONG, 500.0))[0], H(p64.step(big, RA_LONG, 500.0))[0]) < 1e-5


def test_reference_nonlinear_validation_case():
    """Tests/Run_Tests.py:170-196 (thin gap, l=10): d=0.353, Ra=2360, Ra_s=0,
    Tau=1, Pr=1, N_r=24, N_fm=48, dt=0.075, symmetric, t=2000.  The run
    settles on steady convection: the reference's own consistency check
    (inner and outer Nusselt numbers within 0.1%, Main.py:64-66) passes, and
    the mixed-precision run ends within 1e-6 of the float64 kinetic energy."""
    K, Nr, d, dt = 48, 24, 0.353, 0.075
    nr = Nr - 1
    op = ops.build(K, Nr, d, dt, 1.0, 1.0)
    p64 = dev.Plan(op, symmetric=True)
    p32 = dev.Plan(op, symmetric=True, precision="mixed")
    X0 = np.random.default_rng(3).random(3 * K * nr)
    X0 = 1e-3 * X0 / np.linalg.norm(X0)
    n = int(round(2000 / dt))
    x64 = p64.step(T(X0), 2360.0, 0.0, nsteps=n - 1000)
    ke_a = H(p64.diagnostics(x64))[0][1]
    x64 = p64.step(x64, 2360.0, 0.0, nsteps=1000)
    x32 = p32.step(T(X0), 2360.0, 0.0, nsteps=n)
    d64 = H(p64.diagnostics(x64))[0]
    d32 = H(p64.diagnostics(x32))[0]
    assert d64[1] > 1e-3 and abs(d64[1] / ke_a - 1) < 1e-6   # convecting and steady
    Xf = H(x64)[0]
    N = K * nr
    o = O.build_operators(K, Nr, d, dt, 1.0, 1.0)
    nu = O.nusselt(Xf[N:2 * N], d, o.R, o.D, K, nr, check=True)   # raises if inconsistent
    assert abs(nu / d64[2] - 1) < 1e-10
    assert abs(d32[1] / d64[1] - 1) < 1e-6


def test_full_size_properties():
    """Size-independent checks at N_r=40, N_fm=512 (no oracle run): F is
    quadratic and its Frechet derivative satisfies F(X+V)-F(X)-F(V)=DF(X)V."""
    K, Nr = 512, 40
    p = dev.Plan(ops.build(K, Nr, 0.35, 5e-3, 1.0, 1 / 15.0))
    rng = np.random.default_rng(5)
    X = T(1e-2 * rng.standard_normal(3 * K * (Nr - 1)))
    V = T(1e-2 * rng.standard_normal(3 * K * (Nr - 1)))
    F = H(p.nlin_fx(X))
    assert rel(H(p.nlin_fx(2 * X)), 4 * F) < 1e-13
    lhs = H(p.nlin_fx(X + V)) - F - H(p.nlin_fx(V))
    assert rel(H(p.nlin_dfx(V, X)), lhs) < 1e-10
    assert np.abs(H(p.nlin_fx(0 * X))).max() == 0.0


def test_errors_are_raised():
    with pytest.raises(ValueError):
        dev.Plan.build(63, 20, 0.35, 5e-3, 1.0, 1.0)
    p = dev.Plan.build(16, 10, 0.35, 5e-3, 1.0, 1.0)
    with pytest.raises(ValueError):
        p.step(torch.zeros((1, 17), dtype=torch.float64, device="cuda"), 1.0, 1.0)
    with pytest.raises(ValueError):
        p.bsub(torch.zeros((1, 16 * 9), dtype=torch.float64, device="cuda"), 3)


@pytest.mark.parametrize("K,Nr", [(64, 20), (512, 40)])
def test_step_linearized_is_the_central_difference_of_the_step(K, Nr):
    """On the device: (step(X + h v) - step(X - h v)) / (2h) equals
    step_linearized(v, X) up to rounding (the step is quadratic in X), and
    step_linearized agrees with the oracle's."""
    d, dt, Pr, Tau, Ra, Ras = 0.35, 5e-3, 1.0, 1.0 / 15.0, 3000.0, 500.0
    p = dev.Plan.build(K, Nr, d, dt, Pr, Tau, max_batch=2)
    rng = np.random.default_rng(K)
    n3 = 3 * K * (Nr - 1)
    X = 0.3 * rng.standard_normal(n3)
    v = rng.standard_normal(n3)
    lin = H(p.step_linearized(T(v), T(X), Ra, Ras))[0]
    for h in (1.0, 0.1):
        both = H(p.step(T(np.stack([X + h * v, X - h * v])), Ra, Ras))
        cd = (both[0] - both[1]) / (2 * h)
        assert rel(cd, lin) < 1e-11 / h, h
    o = O.build_operators(K, Nr, d, dt, Pr, Tau)
    assert rel(lin, O.step_linearized(v, X, o, Ra, Ras)) < 1e-12
