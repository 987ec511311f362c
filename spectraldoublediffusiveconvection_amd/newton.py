"""Device Newton-Krylov steady solves and pseudo-arclength continuation
(sddc_krylov_*, sddc_gmres, sddc_newton, sddc_arclength_step).

Krylov(plan, restart) owns the Krylov basis of one plan.  ``newton`` runs the
Newton iteration of _Newton (Main.py:430-554) with restarted GMRES on the
device (float64, CGS2 Gram-Schmidt); ``gmres`` solves one linearised system
J x = b with J = PDFX (Main.py:498-521); ``arclength_step`` is one step of
_ContinC (Main.py:754-950) on Y = (X, Ra_T).  States are (3N,) or (1, 3N)
float64 CUDA tensors; Y and its tangent are (3N+1,).

``newton_`` and ``arclength_step_`` work in place on the caller's tensors
and take Ra_T, Ra_S as device tensors: they allocate nothing, which is what
the serving engine's worker threads need.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


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

    def _vec(self, x, extra=0):
        x = x.reshape(-1)
        n = 3 * self.plan.N + extra
        if (x.dtype != torch.float64 or not x.is_cuda or x.numel() != n
                or not x.is_contiguous()):
            raise ValueError("vectors must be contiguous float64 CUDA tensors of length %d" % n)
        return x

    def _scalar(self, v):
        """A device scalar: the caller's (length-1 float64 CUDA tensor) or a new one."""
        if isinstance(v, torch.Tensor):
            if v.dtype != torch.float64 or not v.is_cuda or v.numel() != 1:
                raise ValueError("device parameters must be float64 CUDA tensors of one value")
            return v.reshape(-1)
        return self.plan._params(v, 1)

    @staticmethod
    def _opts(tol_newton=1e-8, tol_gmres=1e-4, rtol=1e-5, atol=0.0, max_restarts=250,
              max_newton=5, fd=False):
        o = _lib.SddcKrylovOpts()
        o.tol_newton, o.tol_gmres, o.rtol, o.atol = tol_newton, tol_gmres, rtol, atol
        o.max_restarts, o.max_newton, o.finite_difference = max_restarts, max_newton, int(bool(fd))
        return o

    def gmres(self, X, b, Ra, Ra_s, x0=None, rtol=1e-5, atol=0.0, max_restarts=250, fd=False):
        """Solve J x = b at state X; returns (x, info dict)."""
        X, b = self._vec(X.contiguous()), self._vec(b.contiguous())
        x = torch.zeros_like(b) if x0 is None else self._vec(x0.contiguous()).clone()
        ra, ras = self._scalar(Ra), self._scalar(Ra_s)
        info = _lib.SddcKrylovInfo()
        o = self._opts(rtol=rtol, atol=atol, max_restarts=max_restarts, fd=fd)
        _lib.check(_lib.lib().sddc_gmres(self._h, _ptr(ra), _ptr(ras), _ptr(X), _ptr(b), _ptr(x),
                                         ctypes.byref(o), ctypes.byref(info), _stream()))
        return x, {"converged": bool(info.converged), "iterations": info.iterations,
                   "restarts": info.restarts, "residual": info.residual, "rhs_norm": info.rhs_norm}

    def newton_(self, X, Ra, Ra_s, tol_newton=1e-8, tol_gmres=1e-4, max_newton=5,
                max_restarts=250, fd=False):
        """Newton iteration on Step(X) - X = 0 in place on X; returns the info dict."""
        X = self._vec(X)
        ra, ras = self._scalar(Ra), self._scalar(Ra_s)
        info = _lib.SddcNewtonInfo()
        o = self._opts(tol_newton=tol_newton, tol_gmres=tol_gmres, max_restarts=max_restarts,
                       max_newton=max_newton, fd=fd)
        _lib.check(_lib.lib().sddc_newton(self._h, _ptr(ra), _ptr(ras), _ptr(X), ctypes.byref(o),
                                          ctypes.byref(info), _stream()))
        return {"converged": bool(info.converged), "iterations": info.iterations,
                "krylov_iterations": info.krylov_iterations, "error": info.error}

    def newton(self, X, Ra, Ra_s, **kw):
        """Newton iteration on Step(X) - X = 0; returns (X_new, info dict)."""
        X = self._vec(X.contiguous()).clone()
        return X, self.newton_(X, Ra, Ra_s, **kw)

    def arclength_step_(self, Y, Y_dot, Ra_s, sign, ds, tol_newton=1e-8, tol_gmres=1e-4,
                        rtol=1e-5, max_restarts=250, fd=False):
        """One pseudo-arclength step in place on Y = (X, Ra_T) and its tangent
        Y_dot (zeros: computed).  Returns (ds_new, info dict); info["status"]
        is one of _lib.ARC_* (the reference raises ValueError for the failures)."""
        Y, Y_dot = self._vec(Y, 1), self._vec(Y_dot, 1)
        ras = self._scalar(Ra_s)
        d = ctypes.c_double(float(ds))
        info = _lib.SddcArclengthInfo()
        o = self._opts(tol_newton=tol_newton, tol_gmres=tol_gmres, rtol=rtol,
                       max_restarts=max_restarts, fd=fd)
        _lib.check(_lib.lib().sddc_arclength_step(
            self._h, _ptr(ras), _ptr(Y), _ptr(Y_dot), float(sign), ctypes.byref(d),
            ctypes.byref(o), ctypes.byref(info), _stream()))
        return d.value, {
            "status": info.status, "iterations": info.iterations, "halvings": info.halvings,
            "krylov_iterations": info.krylov_iterations, "tangent_ok": bool(info.tangent_ok),
            "error_X": info.error_X, "error_mu": info.error_mu}

    def arclength_step(self, Y, Y_dot, Ra_s, sign, ds, **kw):
        """One pseudo-arclength step from Y = (X, Ra_T) with tangent Y_dot
        (zeros: computed).  Returns (Y_new, Y_dot_new, ds_new, info dict)."""
        Y = self._vec(Y.contiguous(), 1).clone()
        Y_dot = self._vec(Y_dot.contiguous(), 1).clone()
        ds, info = self.arclength_step_(Y, Y_dot, Ra_s, sign, ds, **kw)
        return Y, Y_dot, ds, info


class NewtonBatch:
    """Newton-Krylov for up to ``max_batch`` instances of one plan at once
    (sddc_newton_batch_*): _Newton (Main.py:430-554) per instance, the
    instances in lockstep with one batched J v per Arnoldi step and the
    Hessenberg least-squares update on the device.  An instance's result is
    the same bits alone (a batch of one) and batched with others."""

    def __init__(self, plan, restart=100, max_batch=8):
        self.plan = plan
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().sddc_newton_batch_create(plan._h, int(restart), int(max_batch),
                                                       ctypes.byref(h)))
        self._h = h
        plan.max_batch = max(plan.max_batch, int(max_batch))
        self.restart, self.max_batch = int(restart), int(max_batch)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().sddc_newton_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve_(self, X, Ra, Ra_s, tol_newton=1e-8, tol_gmres=1e-4, max_newton=5,
               max_restarts=250, fd=False, check_every=4):
        """Newton on every row of X ((B, 3N) contiguous float64 CUDA tensor,
        updated in place); Ra, Ra_s: device tensors of B values (allocation
        free) or host values.  Returns B info dicts."""
        if (X.dtype != torch.float64 or not X.is_cuda or not X.is_contiguous() or X.dim() != 2
                or X.shape[1] != 3 * self.plan.N):
            raise ValueError("X must be a contiguous (B, 3N) float64 CUDA tensor")
        B = X.shape[0]
        if not 1 <= B <= self.max_batch:
            raise ValueError("batch of %d outside [1, %d]" % (B, self.max_batch))
        ra, ras = self.plan._params(Ra, B), self.plan._params(Ra_s, B)
        o = Krylov._opts(tol_newton=tol_newton, tol_gmres=tol_gmres, max_restarts=max_restarts,
                         max_newton=max_newton, fd=fd)
        o.check_every = int(check_every)
        infos = (_lib.SddcNewtonInfo * B)()
        _lib.check(_lib.lib().sddc_newton_batched(self._h, B, _ptr(ra), _ptr(ras), _ptr(X),
                                                  ctypes.byref(o), infos, _stream()))
        return [{"converged": bool(i.converged), "iterations": i.iterations,
                 "krylov_iterations": i.krylov_iterations, "error": i.error} for i in infos]

    def solve(self, X, Ra, Ra_s, **kw):
        """As solve_ on a copy of X ((B, 3N) or (3N,)); returns (X_new, infos)."""
        X = X.reshape(-1, 3 * self.plan.N).contiguous().clone()
        return X, self.solve_(X, Ra, Ra_s, **kw)
