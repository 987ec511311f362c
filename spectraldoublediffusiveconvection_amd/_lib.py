"""ctypes binding of libsddc.so (include/sddc.h).

The library is built in-tree by ``build_lib.build()``.  There is no CPU
fallback: if the library is missing, every compute call raises.
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsddc.so")

SDDC_OK = 0
SDDC_ERR_INVALID = 1
SDDC_ERR_HIP = 2
SDDC_ERR_UNSUPPORTED = 3
SDDC_ERR_ODD_NFM = 4

SDDC_FP64 = 0
SDDC_MIXED = 1

SDDC_IDCT, SDDC_IDST, SDDC_DCT, SDDC_DST = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)


class SddcOperators(ctypes.Structure):
    _fields_ = [
        ("N_fm", ctypes.c_int), ("N_r", ctypes.c_int),
        ("d", ctypes.c_double), ("dt", ctypes.c_double),
        ("Pr", ctypes.c_double), ("Tau", ctypes.c_double),
        ("D", _dp), ("R", _dp), ("Dr", _dp), ("D2s", _dp), ("D2r2", _dp),
        ("IR2v", _dp), ("IR4v", _dp), ("DT0", _dp), ("Rsq", _dp), ("GR", _dp),
        ("D2_A4", _dp), ("IR2_A4", _dp), ("IR4_A4", _dp),
        ("Linv_A4", _dp), ("Linv_T", _dp), ("Linv_S", _dp),
        ("modes_T", _dp), ("modes_S", _dp),
    ]


class SddcTuning(ctypes.Structure):
    _fields_ = [("chain_segment", ctypes.c_int), ("solve_instances", ctypes.c_int),
                ("sub_batches", ctypes.c_int), ("nab2_solve", ctypes.c_int),
                ("spectral_instances", ctypes.c_int), ("reserved", ctypes.c_int * 3)]


class SddcKrylovOpts(ctypes.Structure):
    _fields_ = [("tol_newton", ctypes.c_double), ("tol_gmres", ctypes.c_double),
                ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
                ("max_restarts", ctypes.c_int), ("max_newton", ctypes.c_int),
                ("finite_difference", ctypes.c_int), ("check_every", ctypes.c_int),
                ("reserved", ctypes.c_int * 4)]


class SddcKrylovInfo(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                ("restarts", ctypes.c_int), ("residual", ctypes.c_double),
                ("rhs_norm", ctypes.c_double)]


class SddcNewtonInfo(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                ("krylov_iterations", ctypes.c_int), ("error", ctypes.c_double)]


class SddcArclengthInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int),
                ("halvings", ctypes.c_int), ("krylov_iterations", ctypes.c_int),
                ("tangent_ok", ctypes.c_int), ("reserved", ctypes.c_int),
                ("error_X", ctypes.c_double), ("error_mu", ctypes.c_double)]


ARC_OK, ARC_PREDICTOR_FAILED, ARC_STEP_TOO_SMALL, ARC_CORRECTOR_FAILED = 0, 1, 2, 3


class SddcError(RuntimeError):
    """A libsddc call failed (HIP error or unsupported size)."""


_lib = None

# name -> (restype, argtypes)
_V = ctypes.c_void_p
_I = ctypes.c_int
_SIGS = {
    "sddc_plan_create": (_I, [ctypes.POINTER(SddcOperators), _I, _I, _I, ctypes.POINTER(_V)]),
    "sddc_plan_create_tuned": (_I, [ctypes.POINTER(SddcOperators), _I, _I, _I,
                                    ctypes.POINTER(SddcTuning), ctypes.POINTER(_V)]),
    "sddc_plan_tuning": (_I, [_V, ctypes.POINTER(SddcTuning)]),
    "sddc_plan_reserve": (_I, [_V, _I]),
    "sddc_plan_destroy": (_I, [_V]),
    "sddc_plan_info": (_I, [_V, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_I)]),
    "sddc_transform": (_I, [_I, _V, _V, _I, _I, _I, _V]),
    "sddc_theta_ops": (_I, [_V, _I, _V, _V, _V]),
    "sddc_derivatives": (_I, [_V, _I, _V, _V, _V]),
    "sddc_nlin_fx": (_I, [_V, _I, _V, _V, _V]),
    "sddc_nlin_dfx": (_I, [_V, _I, _V, _V, _V, _V]),
    "sddc_bsub": (_I, [_V, _I, _I, _V, _V, _V]),
    "sddc_step": (_I, [_V, _I, _V, _V, _V, _V, _I, _I, _V]),
    "sddc_step_static": (_I, [_V, _I, _V, _V, _V, _V, _I, _I, _V]),
    "sddc_step_linearized": (_I, [_V, _I, _V, _V, _V, _V, _V, _V]),
    "sddc_diagnostics": (_I, [_V, _I, _V, _V, _V]),
    "sddc_graph_create": (_I, [_V, _I, _V, _V, _V, _I, _V, ctypes.POINTER(_V)]),
    "sddc_graph_create_multi": (_I, [_I, ctypes.POINTER(_V), ctypes.POINTER(_I),
                                     ctypes.POINTER(_V), ctypes.POINTER(_V), ctypes.POINTER(_V),
                                     _I, _V, ctypes.POINTER(_V)]),
    "sddc_graph_launch": (_I, [_V, _V]),
    "sddc_graph_destroy": (_I, [_V]),
    "sddc_run_stage": (_I, [_V, _I, _I, _V, _V, _V, _V]),
    "sddc_krylov_create": (_I, [_V, _I, ctypes.POINTER(_V)]),
    "sddc_krylov_destroy": (_I, [_V]),
    "sddc_gmres": (_I, [_V, _V, _V, _V, _V, _V, ctypes.POINTER(SddcKrylovOpts),
                        ctypes.POINTER(SddcKrylovInfo), _V]),
    "sddc_newton": (_I, [_V, _V, _V, _V, ctypes.POINTER(SddcKrylovOpts),
                         ctypes.POINTER(SddcNewtonInfo), _V]),
    "sddc_newton_batch_create": (_I, [_V, _I, _I, ctypes.POINTER(_V)]),
    "sddc_newton_batch_destroy": (_I, [_V]),
    "sddc_newton_batched": (_I, [_V, _I, _V, _V, _V, ctypes.POINTER(SddcKrylovOpts),
                               ctypes.POINTER(SddcNewtonInfo), _V]),
    "sddc_arclength_step": (_I, [_V, _V, _V, _V, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(SddcKrylovOpts),
                                 ctypes.POINTER(SddcArclengthInfo), _V]),
    "sddc_last_error": (ctypes.c_char_p, []),
    "sddc_version": (ctypes.c_char_p, []),
}

EXPORTED = tuple(_SIGS)


def lib():
    """Load libsddc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libsddc.so is not built (expected at %s); run "
                "`python -m spectraldoublediffusiveconvection_amd.build_lib`" % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc):
    """Raise the Python exception matching a libsddc status code."""
    if rc == SDDC_OK:
        return
    msg = lib().sddc_last_error().decode(errors="replace")
    if rc in (SDDC_ERR_INVALID, SDDC_ERR_ODD_NFM):
        raise ValueError(msg)
    if rc == SDDC_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise SddcError(msg)
