"""Device Newton-Krylov steady solves and pseudo-arclength continuation
(sddc_krylov_*, sddc_gmres, sddc_newton, sddc_arclength_step).

Krylov(plan, restart) owns the Krylov basis of one plan.  ``newton`` runs the
Newton iteration of _Newton (Main.py:430-554) with restarted GMRES on the
device (float64, CGS2 Gram-Schmidt); ``gmres`` solves one linearised system
J x = b with J = PDFX (Main.py:498-521); ``arclength_step`` is one step of
_ContinC (Main.py:754-950) on Y = (X, Ra_T).  States are (3N,) or (1, 3N)
float64 CUDA tensors; Y and its tangent are (3N+1,).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Krylov:
    def __init__(self, plan, restart=300):
        self.plan = plan
        plan.reserve(1)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().sddc_krylov_create(plan._h, int(restart), ctypes.byref(h)))
        self._h = h
        self.restart = int(restart)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().sddc_krylov_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _vec(self, x):
        x = x.reshape(-1).contiguous()
        if x.dtype != torch.float64 or not x.is_cuda or x.numel() != 3 * self.plan.N:
            raise ValueError("vectors must be float64 CUDA tensors of length 3N")
        return x

    @staticmethod
    def _opts(tol_newton=1e-8, tol_gmres=1e-4, rtol=1e-5, atol=0.0, max_restarts=250,
              max_newton=5, fd=False):
        o = _lib.SddcKrylovOpts()
        o.tol_newton, o.tol_gmres, o.rtol, o.atol = tol_newton, tol_gmres, rtol, atol
        o.max_restarts, o.max_newton, o.finite_difference = max_restarts, max_newton, int(bool(fd))
        return o

    def gmres(self, X, b, Ra, Ra_s, x0=None, rtol=1e-5, atol=0.0, max_restarts=250, fd=False):
        """Solve J x = b at state X; returns (x, info dict)."""
        X, b = self._vec(X), self._vec(b)
        x = torch.zeros_like(b) if x0 is None else self._vec(x0).clone()
        ra, ras = self.plan._params(Ra, 1), self.plan._params(Ra_s, 1)
        info = _lib.SddcKrylovInfo()
        o = self._opts(rtol=rtol, atol=atol, max_restarts=max_restarts, fd=fd)
        _lib.check(_lib.lib().sddc_gmres(self._h, _ptr(ra), _ptr(ras), _ptr(X), _ptr(b), _ptr(x),
                                         ctypes.byref(o), ctypes.byref(info),
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return x, {"converged": bool(info.converged), "iterations": info.iterations,
                   "restarts": info.restarts, "residual": info.residual, "rhs_norm": info.rhs_norm}

    def newton(self, X, Ra, Ra_s, tol_newton=1e-8, tol_gmres=1e-4, max_newton=5,
               max_restarts=250, fd=False):
        """Newton iteration on Step(X) - X = 0; returns (X_new, info dict)."""
        X = self._vec(X).clone()
        ra, ras = self.plan._params(Ra, 1), self.plan._params(Ra_s, 1)
        info = _lib.SddcNewtonInfo()
        o = self._opts(tol_newton=tol_newton, tol_gmres=tol_gmres, max_restarts=max_restarts,
                       max_newton=max_newton, fd=fd)
        _lib.check(_lib.lib().sddc_newton(self._h, _ptr(ra), _ptr(ras), _ptr(X), ctypes.byref(o),
                                          ctypes.byref(info),
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return X, {"converged": bool(info.converged), "iterations": info.iterations,
                   "krylov_iterations": info.krylov_iterations, "error": info.error}

    def arclength_step(self, Y, Y_dot, Ra_s, sign, ds, tol_newton=1e-8, tol_gmres=1e-4,
                       rtol=1e-5, max_restarts=250, fd=False):
        """One pseudo-arclength step from Y = (X, Ra_T) with tangent Y_dot
        (zeros: computed).  Returns (Y_new, Y_dot_new, ds_new, info dict);
        info["status"] is one of _lib.ARC_* (the reference raises ValueError
        for the failures)."""
        n = 3 * self.plan.N
        Y = Y.reshape(-1).contiguous().clone()
        Y_dot = Y_dot.reshape(-1).contiguous().clone()
        for v in (Y, Y_dot):
            if v.dtype != torch.float64 or not v.is_cuda or v.numel() != n + 1:
                raise ValueError("Y and Y_dot must be float64 CUDA tensors of length 3N+1")
        ras = self.plan._params(Ra_s, 1)
        d = ctypes.c_double(float(ds))
        info = _lib.SddcArclengthInfo()
        o = self._opts(tol_newton=tol_newton, tol_gmres=tol_gmres, rtol=rtol,
                       max_restarts=max_restarts, fd=fd)
        _lib.check(_lib.lib().sddc_arclength_step(
            self._h, _ptr(ras), _ptr(Y), _ptr(Y_dot), float(sign), ctypes.byref(d),
            ctypes.byref(o), ctypes.byref(info),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return Y, Y_dot, d.value, {
            "status": info.status, "iterations": info.iterations, "halvings": info.halvings,
            "krylov_iterations": info.krylov_iterations, "tangent_ok": bool(info.tangent_ok),
            "error_X": info.error_X, "error_mu": info.error_mu}
