"""sddc_step_static, the allocation-free step (no counterpart in the
reference; for callers that cannot allocate device memory), against
sddc_step through the C ABI: two calls in a row, the second on a smaller
batch, must give the same bits as sddc_step on the same inputs; a plan whose
reserved batch is 1 steps a larger batch the same way; a batch that does not
fit the static scratch is refused before any kernel runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from spectraldoublediffusiveconvection_amd import _lib  # noqa: E402
from spectraldoublediffusiveconvection_amd import device as dev  # noqa: E402

CASE = dict(d=0.35, dt=5e-3, Pr=1.0, Tau=1.0 / 15.0)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _run(name, plan, X, Ra, Ras, nsteps):
    out = torch.empty_like(X)
    rc = getattr(_lib.lib(), name)(plan._h, X.shape[0], _p(Ra), _p(Ras), _p(X), _p(out),
                                   nsteps, 0, _stream())
    _lib.check(rc)
    torch.cuda.synchronize()
    return out


def _inputs(B, n3, seed):
    X = np.random.default_rng(seed).standard_normal((B, n3))
    X = 1e-3 * X / np.linalg.norm(X, axis=1, keepdims=True)
    Ra = np.linspace(2500.0, 3000.0, B)
    Ras = np.full(B, 500.0)
    return [torch.as_tensor(a, device="cuda") for a in (X, Ra, Ras)]


@pytest.mark.parametrize("K,Nr,precision,sym,B1,B2", [
    (64, 20, "fp64", False, 8, 3),
    (64, 20, "mixed", True, 6, 2),
    (512, 40, "mixed", False, 24, 5),
    (512, 40, "fp64", False, 16, 3),
])
def test_static_step_same_bits_as_step(K, Nr, precision, sym, B1, B2):
    print("case K=%d N_r=%d %s symmetric=%s B=%d then %d" % (K, Nr, precision, sym, B1, B2),
          flush=True)
    plan = dev.Plan.build(K, Nr, **CASE, symmetric=sym, precision=precision, max_batch=B1)
    n3 = 3 * K * (Nr - 1)
    for B, seed in ((B1, 1), (B2, 2)):          # the second call on a smaller batch
        X, Ra, Ras = _inputs(B, n3, seed)
        got = _run("sddc_step_static", plan, X, Ra, Ras, 3)
        want = _run("sddc_step", plan, X, Ra, Ras, 3)
        assert torch.equal(got, want), (B, float((got - want).abs().max()))
    # the reserved batch of the plan does not limit the static step
    small = dev.Plan.build(K, Nr, **CASE, symmetric=sym, precision=precision, max_batch=1,
                           tuning=plan.tuning())
    X, Ra, Ras = _inputs(B1, n3, 3)
    got = _run("sddc_step_static", small, X, Ra, Ras, 2)
    want = _run("sddc_step", plan, X, Ra, Ras, 2)
    assert torch.equal(got, want)


def test_static_step_refuses_a_batch_above_the_scratch():
    K, Nr = 512, 40
    plan = dev.Plan.build(K, Nr, **CASE, precision="mixed", max_batch=1)
    B = 64                                       # about 33 fit in mixed precision
    X = torch.zeros((B, 3 * K * (Nr - 1)), dtype=torch.float64, device="cuda")
    Ra = torch.zeros(B, dtype=torch.float64, device="cuda")
    out = torch.empty_like(X)
    rc = _lib.lib().sddc_step_static(plan._h, B, _p(Ra), _p(Ra), _p(X), _p(out), 1, 0, _stream())
    assert rc == _lib.SDDC_ERR_INVALID
    msg = _lib.lib().sddc_last_error().decode()
    assert "static scratch" in msg and "fit" in msg, msg
    with pytest.raises(ValueError, match="static scratch"):
        plan.step_static(X, Ra, Ra, out)
