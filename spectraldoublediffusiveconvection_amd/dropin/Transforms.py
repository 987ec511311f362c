"""Drop-in for the reference's Transforms.py (grid, DCT, IDCT, DST, IDST and
their scaling helpers), with the transforms computed by libsddc on the GPU.

Signatures, argument meaning and return conventions follow Transforms.py:8-129:
arrays are float64, transforms act along the last axis, results are new arrays.
The scaling helpers act on the last axis whatever `axis` says, so `axis`
other than the last axis raises ValueError.  Shortening the inverse sine
transform below its input length (n < len) has no counterpart in the C ABI
and raises NotImplementedError; the reference never calls it that way.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from ..device import require_gpu

L = np.pi


def grid(N):
    """theta_i = pi (2i+1) / (2N), i < N (Transforms.py:8-13)."""
    dx = L / N
    return np.asarray([dx * (2.0 * i + 1.0) / 2 for i in range(N)])


# scalings (Transforms.py:16-70): host-side relabelling of coefficients
def sinusoid_to_IDCT(f_hat):
    f = np.array(f_hat, dtype=np.float64, copy=True)
    f[..., 1:] *= 0.5
    return f


def DCT_to_sinusoid(f_hat_scaled):
    f = np.array(f_hat_scaled, dtype=np.float64, copy=True)
    f *= 1 / f.shape[-1]
    f[..., 0] *= 0.5
    return f


def sinusoid_to_IDST(g_hat):
    g = np.array(g_hat, dtype=np.float64, copy=True) * 0.5
    g[..., :-1] = g[..., 1:].copy()
    g[..., -1] = 0.0
    return g


def DST_to_sinusoid(g_hat_scaled):
    g = np.array(g_hat_scaled, dtype=np.float64, copy=True)
    g[..., 1:] = g[..., :-1].copy()
    g[..., 0] = 0
    g *= 1 / g.shape[-1]
    return g


def _check_axis(a, axis):
    if axis not in (-1, a.ndim - 1):
        raise ValueError("only the last axis is supported (the reference's scalings act on it)")


def _run(kind, a, n_out):
    """rows of `a` (last axis) through sddc_transform; returns (..., n_out)."""
    require_gpu()
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    lead = a.shape[:-1]
    len_in = a.shape[-1]
    rows = int(np.prod(lead)) if lead else 1
    src = torch.as_tensor(a.reshape(rows, len_in), device="cuda")
    out = torch.empty((rows, n_out), dtype=torch.float64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.lib().sddc_transform(kind, ctypes.c_void_p(src.data_ptr()),
                                         ctypes.c_void_p(out.data_ptr()), rows, len_in, n_out,
                                         stream))
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(lead + (n_out,))


def DST(g, n=None, axis=-1):
    """Sine coefficients of grid values (Transforms.py:73-85); n truncates."""
    g = np.asarray(g, dtype=np.float64)
    _check_axis(g, axis)
    M = g.shape[-1]
    return _run(_lib.SDDC_DST, g, M if n is None else min(int(n), M))


def IDST(g_hat, n=None, axis=-1):
    """Grid values of a sine series (Transforms.py:87-100); n zero-pads."""
    g_hat = np.asarray(g_hat, dtype=np.float64)
    _check_axis(g_hat, axis)
    K = g_hat.shape[-1]
    n = K if n is None else int(n)
    if n < K:
        raise NotImplementedError("IDST with n below the input length")
    return _run(_lib.SDDC_IDST, g_hat, n)


def DCT(f, n=None, axis=-1):
    """Cosine coefficients of grid values (Transforms.py:102-114); n truncates."""
    f = np.asarray(f, dtype=np.float64)
    _check_axis(f, axis)
    M = f.shape[-1]
    return _run(_lib.SDDC_DCT, f, M if n is None else min(int(n), M))


def IDCT(f_hat, n=None, axis=-1):
    """Grid values of a cosine series (Transforms.py:116-129); n zero-pads,
    and n below the input length truncates the series first (as the type-3
    DCT with n < len does)."""
    f_hat = np.asarray(f_hat, dtype=np.float64)
    _check_axis(f_hat, axis)
    K = f_hat.shape[-1]
    n = K if n is None else int(n)
    if n < K:
        f_hat = f_hat[..., :n]
    return _run(_lib.SDDC_IDCT, f_hat, n)
