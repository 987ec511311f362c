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
