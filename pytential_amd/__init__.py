"""MI355X-native QBX layer-potential evaluation (drop-in for the FMM-backed
path of pytential's QBXLayerPotentialSource).

The compute path is libqbx_mi355x.so (hand-written HIP for gfx950, C ABI in
include/qbx_mi355x.h); this package is the host-side mirror of the
reference's Python interface.
"""
from .discretization import (CurveDiscretization, discretize_curve, ellipse,  # noqa: F401
                             grid_targets, starfish)

__version__ = "0.1.0"


def __getattr__(name):
    # GPU-facing modules import torch; load them lazily
    if name in ("QBXLayerPotentialSource", "fmm_level_to_order_for_tolerance"):
        from . import qbx
        return getattr(qbx, name)
    if name in ("bind", "sym"):
        from . import symbolic
        return getattr(symbolic, name)
    raise AttributeError(name)
