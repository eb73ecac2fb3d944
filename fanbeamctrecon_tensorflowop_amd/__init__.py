"""MI355X-native (gfx950) fan-beam CT operators: FP, BP, FBP and FDK backprojection.

Drop-in replacements for the TF1 ops of xtie97/FanBeamCTRecon_TensorflowOp
(``fp_fan_curve``, ``bp_fan_curve``, ``fbp_fan_curve``, ``BackProjection``), backed by
hand-written HIP kernels in ``libfbct.so`` behind the C ABI of ``include/fbct.h``.
"""
from .plan import FanBeamPlan, get_plan
from .ops import (BackProjection, back_project, bp_fan_curve, fbp_fan_curve, fdk, filtered_back_project,
                  forward_project, fp_adjoint, fp_fan_curve)
from . import compat, workloads

__all__ = [
    "FanBeamPlan", "get_plan", "fp_fan_curve", "bp_fan_curve", "fbp_fan_curve", "BackProjection",
    "forward_project", "fp_adjoint", "back_project", "filtered_back_project", "fdk", "compat", "workloads",
]
__version__ = "1.0.0"
