"""Device plan: one size class (N_fm, N_r, d, dt, Pr, Tau) resident on the GPU.

PyTorch provides device memory and the current stream only; every operation
is a libsddc kernel.  States are float64 tensors of shape (B, 3N) in the
reference layout X[f*N + k*nr + i] (Matrix_Operators.py:529-627).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import operators as _ops

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev_ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("libsddc needs a ROCm GPU (torch.cuda.is_available() is False)")


def check_out(out, X):
    """The output tensor of a call on the batch X: a new one when out is None,
    else the caller's, which must be a contiguous float64 tensor on X's device
    with exactly X.numel() values (the kernels write B*3N doubles through its
    data pointer).  Raises ValueError otherwise."""
    if out is None:
        return torch.empty_like(X)
    if not isinstance(out, torch.Tensor) or out.dtype != torch.float64:
        raise ValueError("out must be a float64 tensor")
    if out.device != X.device:
        raise ValueError("out is on %s, the states on %s" % (out.device, X.device))
    if not out.is_contiguous():
        raise ValueError("out must be contiguous")
    if out.numel() != X.numel():
        raise ValueError("out has %d values, the batch needs %d" % (out.numel(), X.numel()))
    return out


class Plan:
    """Operators of one size class uploaded once (Build_Matrix_Operators,
    Main.py:179-225) plus workspace for up to ``max_batch`` instances."""

    def __init__(self, ops: _ops.OperatorSet, symmetric=False, precision="fp64", max_batch=1,
                 tuning=None):
        require_gpu()
        L = _lib.lib()
        self.ops = ops
        self.N_fm, self.N_r, self.nr = ops.N_fm, ops.N_r, ops.N_r - 1
        self.N = self.N_fm * self.nr
        self.symmetric = bool(symmetric)
        self.precision = {"fp64": _lib.SDDC_FP64, "mixed": _lib.SDDC_MIXED}[precision]
        self.precision_name = precision
        keep = {}

        def arr(name):
            a = np.ascontiguousarray(np.asarray(getattr(ops, name), dtype=np.float64))
            keep[name] = a
            return _ptr(a)

        s = _lib.SddcOperators()
        s.N_fm, s.N_r = ops.N_fm, ops.N_r
        s.d, s.dt, s.Pr, s.Tau = float(ops.d), float(ops.dt), float(ops.Pr), float(ops.Tau)
        for name in ("D", "R", "Dr", "D2s", "D2r2", "IR2v", "IR4v", "DT0", "Rsq", "GR", "D2_A4",
                     "IR2_A4", "IR4_A4", "Linv_A4", "Linv_T", "Linv_S"):
            setattr(s, name, arr(name))
        for name in ("modes_T", "modes_S"):   # optional: NULL selects the block solve
            if getattr(ops, name, None) is not None:
                setattr(s, name, arr(name))
        h = ctypes.c_void_p()
        tun = None
        if tuning:
            tun = _lib.SddcTuning()
            tun.chain_segment = int(tuning.get("chain_segment", 0))
            tun.solve_instances = int(tuning.get("solve_instances", 0))
            tun.sub_batches = int(tuning.get("sub_batches", 0))
            tun.nab2_solve = int(tuning.get("nab2_solve", 0))
            tun.spectral_instances = int(tuning.get("spectral_instances", 0))
        _lib.check(L.sddc_plan_create_tuned(ctypes.byref(s), self.precision, int(self.symmetric),
                                            int(max_batch), ctypes.byref(tun) if tun else None,
                                            ctypes.byref(h)))
        self._h = h
        self.max_batch = int(max_batch)
        self.device = torch.device("cuda", torch.cuda.current_device())

    @classmethod
    def build(cls, N_fm, N_r, d, dt, Pr, Tau, symmetric=False, precision="fp64", max_batch=1,
              tuning=None):
        return cls(_ops.build(N_fm, N_r, d, dt, Pr, Tau), symmetric, precision, max_batch, tuning)

    def tuning(self):
        """Launch shapes in use: {'chain_segment': rows, 'solve_instances': n
        (0 = by batch), 'sub_batches': parts of a batch on forked streams in graphs,
        'nab2_solve': 1 per-mode block inverses, 2 eigenbasis,
        'spectral_instances': instances per spectral workgroup (0 = by batch)}."""
        t = _lib.SddcTuning()
        _lib.check(_lib.lib().sddc_plan_tuning(self._h, ctypes.byref(t)))
        return {"chain_segment": t.chain_segment, "solve_instances": t.solve_instances,
                "sub_batches": t.sub_batches, "nab2_solve": t.nab2_solve,
                "spectral_instances": t.spectral_instances}

    def close(self):
        """Free the plan (refused, ValueError, while a graph holds its workspace)."""
        if getattr(self, "_h", None):
            _lib.check(_lib.lib().sddc_plan_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------------------- helpers
    def reserve(self, B):
        if B > self.max_batch:
            _lib.check(_lib.lib().sddc_plan_reserve(self._h, int(B)))
            self.max_batch = int(B)

    def info(self):
        ws = ctypes.c_longlong()
        lds = ctypes.c_int()
        _lib.check(_lib.lib().sddc_plan_info(self._h, ctypes.byref(ws), ctypes.byref(lds)))
        return {"workspace_bytes": ws.value, "lds_nonlinear": lds.value}

    def _state(self, X):
        if not (isinstance(X, torch.Tensor) and X.is_cuda and X.dtype == torch.float64):
            raise TypeError("states must be float64 CUDA tensors")
        if X.device != self.device:
            raise ValueError("states are on %s, the plan on %s" % (X.device, self.device))
        if X.dim() not in (1, 2):
            raise ValueError("states must be (3N,) or (B, 3N)")
        X = X.contiguous()
        if X.dim() == 1:
            X = X.view(1, -1)
        if X.shape[-1] != 3 * self.N:
            raise ValueError("state length %d != 3*N_fm*nr = %d" % (X.shape[-1], 3 * self.N))
        self.reserve(X.shape[0])
        return X

    def _out(self, out, X):
        return check_out(out, X)

    def _params(self, v, B):
        if isinstance(v, torch.Tensor):
            t = v.to(device=self.device, dtype=torch.float64).reshape(-1).contiguous()
            if t.numel() == 1 and B > 1:
                t = t.expand(B).contiguous()
        else:
            arr = np.array(np.broadcast_to(np.asarray(v, dtype=np.float64), (B,)))
            t = torch.as_tensor(arr, device=self.device)
        if t.numel() != B:
            raise ValueError("parameter array length must equal the batch size")
        return t

    # ------------------------------------------------------------------- calls
    def step(self, X, Ra, Ra_s, nsteps=1, linear=False, out=None):
        """nsteps of Step_Python (Main.py:255-283) with the loop's symmetry
        mask (Main.py:323).  X: (B, 3N) float64 CUDA tensor."""
        X = self._state(X)
        B = X.shape[0]
        out = self._out(out, X)
        ra, ras = self._params(Ra, B), self._params(Ra_s, B)
        _lib.check(_lib.lib().sddc_step(self._h, B, _dev_ptr(ra), _dev_ptr(ras), _dev_ptr(X),
                                        _dev_ptr(out), int(nsteps), int(bool(linear)), _stream()))
        return out

    def step_static(self, X, Ra, Ra_s, out, nsteps=1, linear=False):
        """step() without the plan's workspace (sddc_step_static): the
        intermediates live in the library's static per-GPU scratch, so the
        call allocates no device memory and the batch may exceed the plan's
        reserved size.  X, out: contiguous (B, 3N) float64 tensors on the
        plan's device; Ra, Ra_s: device tensors of length B (host values
        would be copied into new tensors).  Calls overlapping on one GPU race."""
        if not (isinstance(X, torch.Tensor) and X.is_cuda and X.dtype == torch.float64
                and X.is_contiguous() and X.device == self.device):
            raise ValueError("X must be a contiguous float64 tensor on the plan's device")
        X = X.view(1, -1) if X.dim() == 1 else X
        if X.dim() != 2 or X.shape[-1] != 3 * self.N:
            raise ValueError("states must be (B, 3N) with 3N = %d" % (3 * self.N))
        B = X.shape[0]
        out = check_out(out, X)
        ra, ras = self._params(Ra, B), self._params(Ra_s, B)
        _lib.check(_lib.lib().sddc_step_static(self._h, B, _dev_ptr(ra), _dev_ptr(ras),
                                               _dev_ptr(X), _dev_ptr(out), int(nsteps),
                                               int(bool(linear)), _stream()))
        return out

    def step_linearized(self, dv, X, Ra, Ra_s, out=None):
        """PDFX without '- dv' (Main.py:498-521)."""
        X = self._state(X)
        dv = self._state(dv)
        B = X.shape[0]
        if dv.shape != X.shape:
            raise ValueError("dv and X must have the same shape")
        out = self._out(out, X)
        ra, ras = self._params(Ra, B), self._params(Ra_s, B)
        _lib.check(_lib.lib().sddc_step_linearized(self._h, B, _dev_ptr(ra), _dev_ptr(ras),
                                                   _dev_ptr(dv), _dev_ptr(X), _dev_ptr(out),
                                                   _stream()))
        return out

    def nlin_fx(self, X):
        X = self._state(X)
        out = torch.empty_like(X)
        _lib.check(_lib.lib().sddc_nlin_fx(self._h, X.shape[0], _dev_ptr(X), _dev_ptr(out), _stream()))
        return out

    def nlin_dfx(self, dv, X):
        X = self._state(X)
        dv = self._state(dv)
        out = torch.empty_like(X)
        _lib.check(_lib.lib().sddc_nlin_dfx(self._h, X.shape[0], _dev_ptr(dv), _dev_ptr(X),
                                            _dev_ptr(out), _stream()))
        return out

    def theta_ops(self, X):
        """(B, 4, N): J_theta_RT, DT0_theta, A2_SINE, A2_SINE_R2 of psi."""
        X = self._state(X)
        out = torch.empty((X.shape[0], 4, self.N), dtype=torch.float64, device=X.device)
        _lib.check(_lib.lib().sddc_theta_ops(self._h, X.shape[0], _dev_ptr(X), _dev_ptr(out), _stream()))
        return out

    def derivatives(self, X):
        """(B, 9, nr, N_fm) slabs of Derivatives (Matrix_Operators.py:630-740)."""
        X = self._state(X)
        out = torch.empty((X.shape[0], 9, self.nr, self.N_fm), dtype=torch.float64, device=X.device)
        _lib.check(_lib.lib().sddc_derivatives(self._h, X.shape[0], _dev_ptr(X), _dev_ptr(out), _stream()))
        return out

    def bsub(self, g, field):
        """Back substitution of one field: 0 A4 (psi), 1 NAB2 with Linv_T, 2 with Linv_S."""
        g = g.contiguous()
        if g.dim() == 1:
            g = g.view(1, -1)
        if g.shape[-1] != self.N:
            raise ValueError("right-hand side length must be N_fm*nr")
        self.reserve(g.shape[0])
        out = torch.empty_like(g)
        _lib.check(_lib.lib().sddc_bsub(self._h, g.shape[0], int(field), _dev_ptr(g), _dev_ptr(out),
                                        _stream()))
        return out

    def diagnostics(self, X):
        """(B, 6): ||X||_2, kinetic energy, Nu_T - 1 and Nu_S - 1 at R_1 (the
        values of Main.py:292-295), Nu_T - 1 and Nu_S - 1 at R_2."""
        X = self._state(X)
        out = torch.empty((X.shape[0], 6), dtype=torch.float64, device=X.device)
        _lib.check(_lib.lib().sddc_diagnostics(self._h, X.shape[0], _dev_ptr(X), _dev_ptr(out), _stream()))
        return out

    def run_stage(self, X, Ra, Ra_s, stage):
        """Launch one stage of the step on the workspace (timing/profiling)."""
        X = self._state(X)
        B = X.shape[0]
        ra, ras = self._params(Ra, B), self._params(Ra_s, B)
        _lib.check(_lib.lib().sddc_run_stage(self._h, B, int(stage), _dev_ptr(X), _dev_ptr(ra),
                                             _dev_ptr(ras), _stream()))

    def graph(self, X, Ra, Ra_s, nsteps=1):
        """Capture nsteps in-place steps of the batch X into a HIP graph."""
        return StepGraph(self, X, Ra, Ra_s, nsteps)


class StepGraph:
    """HIP graph of nsteps in-place steps of a fixed batch (sddc_graph_*)."""

    def __init__(self, plan, X, Ra, Ra_s, nsteps):
        self.plan = plan
        if not (isinstance(X, torch.Tensor) and X.is_contiguous()):
            # the graph steps X in place through its data pointer; a
            # contiguous copy would be stepped instead of the caller's tensor
            raise ValueError("a graph's states must be a contiguous tensor")
        self.X = plan._state(X)
        B = self.X.shape[0]
        self.ra, self.ras = plan._params(Ra, B), plan._params(Ra_s, B)
        # capture and replay on a dedicated stream (default priority), ordered
        # against the caller's stream with events
        self.stream = torch.cuda.Stream()
        self.stream.wait_stream(torch.cuda.current_stream())
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().sddc_graph_create(plan._h, B, _dev_ptr(self.ra), _dev_ptr(self.ras),
                                                _dev_ptr(self.X), int(nsteps),
                                                ctypes.c_void_p(self.stream.cuda_stream),
                                                ctypes.byref(h)))
        self._h = h

    def launch(self):
        cur = torch.cuda.current_stream()
        self.stream.wait_stream(cur)
        _lib.check(_lib.lib().sddc_graph_launch(self._h, ctypes.c_void_p(self.stream.cuda_stream)))
        cur.wait_stream(self.stream)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().sddc_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
