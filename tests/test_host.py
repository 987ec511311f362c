"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point of include/sddc.h, the host operator builder produces the
reference's matrices, and argument errors match the reference's."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import sddc_oracle as O
from spectraldoublediffusiveconvection_amd import _lib, operators

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "sddc.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sddc_\w+)\s*\(", text, re.M)))


def test_library_exports_every_header_symbol():
    names = header_functions()
    assert len(names) >= 18
    handle = ctypes.CDLL(_lib.LIB_PATH)   # loads without touching the GPU
    for n in names:
        assert hasattr(handle, n), n
    assert set(names) == set(_lib.EXPORTED)


def test_binding_loads_and_reports_version():
    assert _lib.lib().sddc_version().decode().startswith("sddc")


def test_plan_create_rejects_odd_nfm_before_touching_the_gpu():
    s = _lib.SddcOperators()
    s.N_fm, s.N_r = 63, 10
    h = ctypes.c_void_p()
    rc = _lib.lib().sddc_plan_create(ctypes.byref(s), 0, 0, 1, ctypes.byref(h))
    assert rc == _lib.SDDC_ERR_ODD_NFM
    with pytest.raises(ValueError, match="not even 63"):
        _lib.check(rc)


def test_host_builder_odd_nfm_raises():
    with pytest.raises(ValueError):
        operators.build(63, 10, 0.35, 5e-3, 1.0, 1.0)


@pytest.mark.parametrize("K,Nr", [(64, 20), (16, 12)])
def test_host_operators_equal_reference_construction(K, Nr):
    """operators.build reproduces Build_Matrix_Operators' matrices bit for bit
    (same NumPy operations in the same order as the oracle's restatement)."""
    d, dt, Pr, Tau = 0.35, 5e-3, 1.0, 1 / 15.0
    a = operators.build(K, Nr, d, dt, Pr, Tau)
    o = O.build_operators(K, Nr, d, dt, Pr, Tau)
    pairs = [("D", "D"), ("R", "R"), ("Dr", "Dr"), ("D2s", "D2s"), ("D2r2", "D2r2"),
             ("IR2v", "IR2v"), ("IR4v", "IR4v"), ("DT0", "DT0"), ("Rsq", "Rsq"),
             ("D2_A4", "D2_A4"), ("Linv_A4", "Linv_A4"), ("Linv_T", "Linv_T"),
             ("Linv_S", "Linv_S")]
    for x, y in pairs:
        assert np.array_equal(getattr(a, x), getattr(o, y)), x
    assert np.array_equal(a.IR2_A4, np.diag(o.IR2_A4))
    assert np.array_equal(a.IR4_A4, np.diag(o.IR4_A4))
    assert np.array_equal(-(np.arange(K)[:, None] + 1) * a.GR[None, :], o.grk)


def test_build_is_importable_without_gpu():
    import importlib
    m = importlib.import_module("spectraldoublediffusiveconvection_amd")
    assert os.path.exists(m.LIB_PATH)


def test_output_tensor_checks():
    """Plan.step / step_linearized write B*3N doubles through out's data
    pointer: a wrong dtype, size, device or a strided out is refused before any
    kernel runs (device.check_out, host logic only)."""
    torch = pytest.importorskip("torch")
    from spectraldoublediffusiveconvection_amd import device as dev
    X = torch.zeros((2, 12), dtype=torch.float64)
    assert dev.check_out(None, X).shape == X.shape
    ok = torch.empty((2, 12), dtype=torch.float64)
    assert dev.check_out(ok, X) is ok
    assert dev.check_out(torch.empty(24, dtype=torch.float64), X).numel() == 24
    bad = {
        "float32": torch.empty((2, 12), dtype=torch.float32),
        "smaller": torch.empty((1, 12), dtype=torch.float64),
        "strided": torch.empty((12, 2), dtype=torch.float64).t(),
        "meta": torch.empty((2, 12), dtype=torch.float64, device="meta"),
        "not a tensor": np.zeros((2, 12)),
    }
    for name, out in bad.items():
        with pytest.raises(ValueError):
            dev.check_out(out, X)
            pytest.fail(name)


def _modal_nab2(g, modes, K, nr, h, symmetric=False):
    """The NAB2 chain solved as nr scalar recurrences in the eigenbasis of
    P = r^2 - h r^2 Lap (the form of the device kernel k_nab2_modal)."""
    lam = modes[:nr]
    V = modes[nr:nr + nr * nr].reshape(nr, nr)
    Vi = modes[nr + nr * nr:].reshape(nr, nr)
    gt = g.reshape(K, nr) @ Vi.T
    ft = np.zeros_like(gt)
    for chain in ((1,) if symmetric else (0, 1)):
        bt = np.zeros(nr)
        for j in range(K - 1 - chain, -1, -2):
            q = 1.0 / (lam + h * j * (j + 1))
            ft[j] = (gt[j] - (0.5 if j == 0 else 1.0) * bt) * q
            bt = (lam + h * j * (j - 1)) * q * bt + 2.0 * h * j * q * gt[j]
    return (ft @ V.T).ravel()


@pytest.mark.parametrize("K,Nr,d,dt,Tau", [(64, 20, 0.35, 5e-3, 1 / 15.0), (32, 16, 2.0, 1.0, 1.0),
                                           (48, 30, 0.353, 1.0, 1.0), (512, 40, 0.35, 5e-3, 1 / 15.0)])
def test_nab2_eigenbasis_reproduces_the_reference_blocks(K, Nr, d, dt, Tau):
    """operators.nab2_modes against the reference's per-mode inverses
    (NAB2_TSTEP_MATS, built by the oracle): V diag(1/(lam + h j(j+1))) V^-1
    equals Linv of every mode, and the scalar-recurrence chain equals the
    oracle's NAB2_BSub_TSTEP_V2 for both fields, with and without the
    symmetry mask.  A wrong shift, sign or index of the recurrence fails the
    second check by orders of magnitude more than the tolerance."""
    o = O.build_operators(K, Nr, d, dt, 1.0, Tau)
    ops = operators.build(K, Nr, d, dt, 1.0, Tau)
    nr = Nr - 1
    rng = np.random.default_rng(K + Nr)
    for modes, Linv, h in ((ops.modes_T, o.Linv_T, dt), (ops.modes_S, o.Linv_S, Tau * dt)):
        assert modes is not None and modes.shape == (nr + 2 * nr * nr,)
        lam = modes[:nr]
        V = modes[nr:nr + nr * nr].reshape(nr, nr)
        Vi = modes[nr + nr * nr:].reshape(nr, nr)
        assert np.all(lam > 0)
        for jj in (0, 1, K // 2, K - 2, K - 1):
            j = K - 1 - jj
            Lj = (V * (1.0 / (lam + h * j * (j + 1)))) @ Vi
            assert np.linalg.norm(Lj - Linv[jj]) / np.linalg.norm(Linv[jj]) < 1e-13, jj
        for sym in (False, True):
            g = rng.standard_normal(K * nr)
            ref = O.nab2_bsub(g, Linv, K, nr, h, sym)
            got = _modal_nab2(g, modes, K, nr, h, sym)
            assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 5e-14, sym
