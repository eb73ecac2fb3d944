"""Drop-in operators with the reference op signatures, differentiable through
hand-written matched adjoints.

Reference ops (TF1 custom ops, GPU kernels only):

=====================  ===================================  ==========================
python name            registration                         here
=====================  ===================================  ==========================
``fp_fan_curve``       ``FPFanCurve``  FP/transform.cc:17   :func:`fp_fan_curve`
``bp_fan_curve``       ``BPFanCurve``  BP/transform.cc:17   :func:`bp_fan_curve`
``fbp_fan_curve``      ``FBPFanCurve`` FBP/transform.cc:17  :func:`fbp_fan_curve`
``BackProjection``     ``BackProjection`` FDK/FBP.cc:16     :func:`BackProjection`
=====================  ===================================  ==========================

Each takes ``(volume|projection, angle, parameter, vol_shape, proj_shape)`` exactly as
the reference is called (utils.py:25,35,45; FDK/FDK_test_training.py:111) and returns a
tensor of the attribute-declared output shape (FP/transform.cc:29-33).  They dispatch to
the custom ops ``torch.ops.fbct.*`` registered in ``torch_ops``.  Unlike the
reference, which trusts ``parameter[11]`` for the slice count (FP/transform.cc:75), the
input's element count is checked against it.

Gradients: FP's backward is the exact transpose of its interpolation weights
(``fbct_fp_adjoint``); FBP's backward is filter^T o weighted-BP^T (``fbct_fbp_adjoint``);
BP's backward is BP^T (``fbct_bp_adjoint``).  The operators are linear, so nothing from
the forward pass is kept for the backward pass.
"""
from __future__ import annotations

import math

import torch

from . import torch_ops  # noqa: F401  (registers torch.ops.fbct.*)
from .plan import FanBeamPlan

__all__ = [
    "fp_fan_curve", "bp_fan_curve", "fbp_fan_curve", "BackProjection",
    "forward_project", "back_project", "filtered_back_project", "fdk",
    "FPFunction", "BPFunction", "FBPFunction",
]


class FPFunction(torch.autograd.Function):
    """FP with its matched adjoint as backward.  A bfloat16 image is read as BF16
    storage (FP32 arithmetic); the sinogram is FP32 and the gradient takes the
    image's dtype."""

    @staticmethod
    def forward(ctx, image, plan: FanBeamPlan):
        ctx.plan = plan
        ctx.image_dtype = image.dtype
        return plan.fp(image)

    @staticmethod
    def backward(ctx, grad):
        return ctx.plan.fp_adjoint(grad.contiguous(), out_dtype=ctx.image_dtype), None


class FPAdjointFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, plan: FanBeamPlan):
        ctx.plan = plan
        return plan.fp_adjoint(sino)

    @staticmethod
    def backward(ctx, grad):
        return ctx.plan.fp(grad.contiguous()), None


class BPFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, plan: FanBeamPlan):
        ctx.plan = plan
        return plan.bp(sino)

    @staticmethod
    def backward(ctx, grad):
        return ctx.plan.bp_adjoint(grad.contiguous()), None


class FBPFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, plan: FanBeamPlan):
        ctx.plan = plan
        return plan.fbp(sino)

    @staticmethod
    def backward(ctx, grad):
        return ctx.plan.fbp_adjoint(grad.contiguous()), None


def forward_project(image: torch.Tensor, plan: FanBeamPlan) -> torch.Tensor:
    """Differentiable FP: image [n, nY, nX] -> sinogram [n, nViews, nCols]."""
    return FPFunction.apply(image, plan)


def fp_adjoint(sino: torch.Tensor, plan: FanBeamPlan) -> torch.Tensor:
    """Differentiable exact transpose of :func:`forward_project`."""
    return FPAdjointFunction.apply(sino, plan)


def back_project(sino: torch.Tensor, plan: FanBeamPlan) -> torch.Tensor:
    return BPFunction.apply(sino, plan)


def filtered_back_project(sino: torch.Tensor, plan: FanBeamPlan) -> torch.Tensor:
    return FBPFunction.apply(sino, plan)


def fdk(proj: torch.Tensor, plan: FanBeamPlan) -> torch.Tensor:
    """Reference FDK op semantics (nZ = nRows = 1): the FBP of each slice."""
    return FBPFunction.apply(proj, plan)


# ------------------------------------------------------------------ drop-ins

def _attr_list(a):
    """An op attribute (tensor, array or sequence of floats) as a list of Python floats."""
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().reshape(-1).tolist()
    elif hasattr(a, "tolist"):
        a = a.reshape(-1).tolist()
    return [float(v) for v in a]


def _shape_list(s):
    return [int(v) for v in s]


def _call(op, x, angle, parameter, vol_shape, proj_shape):
    return op(x, _attr_list(angle), _attr_list(parameter), _shape_list(vol_shape), _shape_list(proj_shape))


def fp_fan_curve(volume, angle, parameter, vol_shape, proj_shape):
    """``FPFanCurve`` (FP/transform.cc:17-37,93-117): volume -> projection
    (``torch.ops.fbct.fp_fan_curve``; backward = matched adjoint)."""
    return _call(torch.ops.fbct.fp_fan_curve, volume, angle, parameter, vol_shape, proj_shape)


def bp_fan_curve(projection, angle, parameter, vol_shape, proj_shape):
    """``BPFanCurve`` (BP/transform.cc:17-36,92-116): projection -> volume (unweighted BP)."""
    return _call(torch.ops.fbct.bp_fan_curve, projection, angle, parameter, vol_shape, proj_shape)


def fbp_fan_curve(projection, angle, parameter, vol_shape, proj_shape):
    """``FBPFanCurve`` (FBP/transform.cc:17-36,92-116): ramp filter + 1/L^2-weighted BP."""
    return _call(torch.ops.fbct.fbp_fan_curve, projection, angle, parameter, vol_shape, proj_shape)


def BackProjection(projection, angle, parameter, vol_shape, proj_shape):  # noqa: N802 (reference name)
    """The FDK op ``BackProjection`` (FDK/FBP.cc:16-35,125-152).

    The reference decodes nZ = nRows = 1 (FDK/FBP.cc:85,98), ignores parameter[11] and
    processes the first slice only (FDK/GPURecon_GPU.cu:178-179).  Here the slice count
    is taken from the input (all slices are reconstructed); with one slice the result
    is the reference's.
    """
    return _call(torch.ops.fbct.back_projection, projection, angle, parameter, vol_shape, proj_shape)
