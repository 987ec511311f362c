"""Plans behind the drop-in functions.

The reference's functions receive matrices (D, R, L_inv, ...) rather than a
handle.  Each distinct operator set gets one device plan, keyed by a hash of
the matrices it was given, and kept in a small LRU cache.  Every compute call
copies its numpy inputs to the GPU and its result back, which is what the
drop-in contract needs (new arrays returned, inputs untouched); code that
owns device tensors calls ``device.Plan`` directly.
"""
from __future__ import annotations

import collections

import numpy as np
import xxhash

from .. import device as _dev
from .. import operators as _ops

_MAX_PLANS = 8
_plans: "collections.OrderedDict[tuple, _dev.Plan]" = collections.OrderedDict()
# Chebyshev geometries seen so far, by the hash of their interior block Dr:
# Derivatives() receives Dr only and recovers D and R here.
_geometries: dict = {}


def fingerprint(*arrays) -> int:
    """64-bit hash of the bytes (and shapes) of arrays or lists of arrays."""
    h = xxhash.xxh3_64()
    for a in arrays:
        items = a if isinstance(a, (list, tuple)) else [a]
        for m in items:
            m = np.ascontiguousarray(np.asarray(m, dtype=np.float64))
            h.update(np.asarray(m.shape, dtype=np.int64).tobytes())
            h.update(m.tobytes())
    return h.intdigest()


def remember_geometry(D, R):
    D = np.asarray(D, dtype=np.float64)
    _geometries[fingerprint(D[1:-1, 1:-1])] = (np.array(D), np.array(R, dtype=np.float64))


def geometry_of_dr(Dr):
    g = _geometries.get(fingerprint(Dr))
    if g is None:
        raise ValueError("Derivatives: Dr does not belong to a Chebyshev geometry built by "
                         "cheb_radial or Build_Matrix_Operators in this process")
    return g


def _get(key, make):
    p = _plans.get(key)
    if p is not None:
        _plans.move_to_end(key)
        return p
    p = make()
    _plans[key] = p
    while len(_plans) > _MAX_PLANS:
        _, old = _plans.popitem(last=False)
        _release(old)
        old.close()
    return p


def _release(plan):
    """Close the Krylov basis of the drop-in Newton solver if it uses plan."""
    from . import Main
    kr = Main._krylov.get("k")
    if kr is not None and kr.plan is plan:
        kr.close()
        Main._krylov.clear()


def _dummy_geometry(nr):
    return _ops.cheb_radial(nr + 1, 1.0)


def geometry_plan(D, R, N_fm, symmetric, d=None):
    """Plan for the operators that depend on (D, R) only (transforms, theta
    recurrences, nonlinear terms, diagnostics)."""
    D = np.asarray(D, dtype=np.float64)
    R = np.asarray(R, dtype=np.float64)
    key = ("geo", int(N_fm), fingerprint(D, R), None if d is None else float(d), bool(symmetric))

    def make():
        remember_geometry(D, R)
        return _dev.Plan(_ops.from_matrices(D, R, int(N_fm), d=d), symmetric=bool(symmetric))
    return _get(key, make)


def theta_plan(N_fm, nr, symmetric, dT0=None):
    """Plan for J_theta_RT / DT0_theta, which need no radial operator except
    the DT0 vector."""
    key = ("theta", int(N_fm), int(nr), None if dT0 is None else fingerprint(dT0), bool(symmetric))

    def make():
        D, R = _dummy_geometry(nr)
        ops = _ops.from_matrices(D, R, int(N_fm))
        if dT0 is not None:
            ops.DT0 = np.ascontiguousarray(np.asarray(dT0, dtype=np.float64))
        return _dev.Plan(ops, symmetric=bool(symmetric))
    return _get(key, make)


def nab2_plan(L_inv, N_fm, nr, dt, symmetric):
    """Plan whose temperature solve uses the caller's block inverses and dt."""
    key = ("nab2", int(N_fm), int(nr), float(dt), fingerprint(L_inv), bool(symmetric))

    def make():
        D, R = _dummy_geometry(nr)
        ops = _ops.from_matrices(D, R, int(N_fm), dt=float(dt), Linv_T=L_inv)
        return _dev.Plan(ops, symmetric=bool(symmetric))
    return _get(key, make)


def a4_plan(L_inv, D2, IR4, IR2, N_fm, nr, dt, symmetric):
    """Plan whose psi solve uses the caller's inverses, D2, IR4, IR2 and dt
    (dt already multiplied by Pr, as the reference passes it)."""
    key = ("a4", int(N_fm), int(nr), float(dt), fingerprint(L_inv, D2, IR4, IR2), bool(symmetric))

    def make():
        D, R = _dummy_geometry(nr)
        ops = _ops.from_matrices(D, R, int(N_fm), dt=float(dt), Pr=1.0, Linv_A4=L_inv,
                                 D2_A4=D2, IR2_A4=IR2, IR4_A4=IR4)
        return _dev.Plan(ops, symmetric=bool(symmetric))
    return _get(key, make)


def step_plan(N_fm, N_r, d, dt, Pr, Tau, symmetric):
    """Plan of Build_Matrix_Operators(N_fm, N_r, d, dt, Pr, Tau)."""
    key = ("step", int(N_fm), int(N_r), float(d), float(dt), float(Pr), float(Tau), bool(symmetric))

    def make():
        ops = _ops.build(int(N_fm), int(N_r), float(d), float(dt), float(Pr), float(Tau))
        remember_geometry(ops.D, ops.R)
        return _dev.Plan(ops, symmetric=bool(symmetric))
    return _get(key, make)


def clear():
    while _plans:
        _, p = _plans.popitem()
        _release(p)
        p.close()
