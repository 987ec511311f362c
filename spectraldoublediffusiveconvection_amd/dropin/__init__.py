"""Drop-in modules with the reference's names: Transforms, Matrix_Operators
and Main, computed by libsddc on the GPU.

    from spectraldoublediffusiveconvection_amd import dropin
    dropin.install()          # registers the three modules under their names
    from Main import _Time_Step, _Newton, Kinetic_Energy, Nusselt

After install(), the reference's own drivers (Time_Cotinuation.py,
Gap_Continuation.py, Run_Tests.py), which import these modules by name, run
on the GPU.
"""
from __future__ import annotations

import sys

NAMES = ("Transforms", "Matrix_Operators", "Main")


def install():
    """Register the drop-in modules in sys.modules under the reference's names."""
    from . import Main, Matrix_Operators, Transforms
    mods = {"Transforms": Transforms, "Matrix_Operators": Matrix_Operators, "Main": Main}
    for name in NAMES:
        sys.modules[name] = mods[name]
    return mods


def uninstall():
    from . import Main, Matrix_Operators, Transforms
    for name, mod in (("Transforms", Transforms), ("Matrix_Operators", Matrix_Operators),
                      ("Main", Main)):
        if sys.modules.get(name) is mod:
            del sys.modules[name]
