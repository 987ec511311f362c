"""MI355X-native (gfx950) pseudospectral double-diffusive convection stepper.

Host code mirrors the reference's Python call surface (Transforms,
Matrix_Operators, Main); every per-step operation runs in hand-written HIP
kernels (libsddc.so) through a C ABI.  PyTorch-ROCm supplies device memory
and streams.
"""
from ._lib import LIB_PATH, SddcError, lib  # noqa: F401

__version__ = "0.1.0"


def device_plan(*args, **kwargs):
    """Plan.build(...) -- see device.Plan."""
    from .device import Plan
    return Plan.build(*args, **kwargs)
