"""The reference operators registered as PyTorch custom ops (``torch.ops.fbct.*``).

Reference registrations (TF1, GPU kernels only):

=================================  ====================================  =========================
reference                          here                                  backward
=================================  ====================================  =========================
``FPFanCurve``  FP/transform.cc:17  ``fbct::fp_fan_curve``                ``fp_fan_curve_adjoint``
``BPFanCurve``  BP/transform.cc:17  ``fbct::bp_fan_curve``                ``bp_fan_curve_adjoint``
``FBPFanCurve`` FBP/transform.cc:17 ``fbct::fbp_fan_curve``               ``fbp_fan_curve_adjoint``
``BackProjection`` FDK/FBP.cc:16    ``fbct::back_projection``             ``fbp_fan_curve_adjoint``
(none)                              ``fbct::*_adjoint`` (the transposes)  the forward op
=================================  ====================================  =========================

Schema of every op: ``(Tensor x, float[] angle, float[] parameter, int[] vol_shape,
int[] proj_shape) -> Tensor`` -- the reference's input tensor plus its four attributes
(FP/transform.cc:18-23; ``angle`` in radians, ``parameter`` the 16-float vector).  The
output has the attribute-declared shape (FP/transform.cc:29-33).  Each op has a fake
(meta) implementation for shape propagation and an autograd formula given by the
matched adjoint op; the operators are linear, so backward keeps nothing from forward.
"""
from __future__ import annotations

import math
from typing import List

import torch

from .plan import get_plan

__all__ = ["OPS"]


def _plan(angle, parameter):
    return get_plan(list(parameter), list(angle))


def _slices(x, plan, per_slice, what):
    n = plan.n_samp
    if n <= 0:
        raise ValueError("%s: parameter[11] (slice count) must be positive, got %d" % (what, n))
    if x.numel() != n * per_slice:
        raise ValueError("%s: input has %d elements but parameter[11]=%d slices need %d"
                         % (what, x.numel(), n, n * per_slice))
    return n


def _shaped(y, shape, what):
    shape = [int(s) for s in shape]
    if math.prod(shape) != y.numel():
        raise ValueError("%s: declared output shape %s holds %d elements, result has %d"
                         % (what, shape, math.prod(shape), y.numel()))
    return y.reshape(shape)


def _img_in(x, plan, what):
    n = _slices(x, plan, plan.nx * plan.ny, what)
    return x.reshape(n, plan.ny, plan.nx)


def _sino_in(x, plan, what):
    n = _slices(x, plan, plan.n_views * plan.n_cols, what)
    return x.reshape(n, plan.n_views, plan.n_cols)


def _fdk_in(x, plan, what):
    per = plan.n_views * plan.n_cols
    if x.numel() % per != 0:
        raise ValueError("%s: input size %d is not a multiple of nViews*nCols=%d" % (what, x.numel(), per))
    return x.reshape(x.numel() // per, plan.n_views, plan.n_cols)


# ------------------------------------------------------------------ forward ops

@torch.library.custom_op("fbct::fp_fan_curve", mutates_args=())
def fp_fan_curve(volume: torch.Tensor, angle: List[float], parameter: List[float], vol_shape: List[int],
                 proj_shape: List[int]) -> torch.Tensor:
    plan = _plan(angle, parameter)
    return _shaped(plan.fp(_img_in(volume, plan, "fp_fan_curve")), proj_shape, "fp_fan_curve")


@torch.library.custom_op("fbct::fp_fan_curve_adjoint", mutates_args=())
def fp_fan_curve_adjoint(projection: torch.Tensor, angle: List[float], parameter: List[float],
                         vol_shape: List[int], proj_shape: List[int]) -> torch.Tensor:
    plan = _plan(angle, parameter)
    return _shaped(plan.fp_adjoint(_sino_in(projection, plan, "fp_fan_curve_adjoint")), vol_shape,
                   "fp_fan_curve_adjoint")


@torch.library.custom_op("fbct::bp_fan_curve", mutates_args=())
def bp_fan_curve(projection: torch.Tensor, angle: List[float], parameter: List[float], vol_shape: List[int],
                 proj_shape: List[int]) -> torch.Tensor:
    plan = _plan(angle, parameter)
    return _shaped(plan.bp(_sino_in(projection, plan, "bp_fan_curve")), vol_shape, "bp_fan_curve")


@torch.library.custom_op("fbct::bp_fan_curve_adjoint", mutates_args=())
def bp_fan_curve_adjoint(volume: torch.Tensor, angle: List[float], parameter: List[float],
                         vol_shape: List[int], proj_shape: List[int]) -> torch.Tensor:
    plan = _plan(angle, parameter)
    return _shaped(plan.bp_adjoint(_img_in(volume, plan, "bp_fan_curve_adjoint")), proj_shape,
                   "bp_fan_curve_adjoint")


@torch.library.custom_op("fbct::fbp_fan_curve", mutates_args=())
def fbp_fan_curve(projection: torch.Tensor, angle: List[float], parameter: List[float], vol_shape: List[int],
                  proj_shape: List[int]) -> torch.Tensor:
    plan = _plan(angle, parameter)
    return _shaped(plan.fbp(_sino_in(projection, plan, "fbp_fan_curve")), vol_shape, "fbp_fan_curve")


@torch.library.custom_op("fbct::fbp_fan_curve_adjoint", mutates_args=())
def fbp_fan_curve_adjoint(volume: torch.Tensor, angle: List[float], parameter: List[float],
                          vol_shape: List[int], proj_shape: List[int]) -> torch.Tensor:
    plan = _plan(angle, parameter)
    n = volume.numel() // (plan.nx * plan.ny)
    if n * plan.nx * plan.ny != volume.numel():
        raise ValueError("fbp_fan_curve_adjoint: input size %d is not a multiple of nX*nY" % volume.numel())
    y = plan.fbp_adjoint(volume.reshape(n, plan.ny, plan.nx))
    # forward shape of the op whose gradient this is (FBP or the FDK op, which takes its
    # slice count from the input rather than from parameter[11])
    return _shaped(y, proj_shape if math.prod(proj_shape) == y.numel() else list(y.shape),
                   "fbp_fan_curve_adjoint")


@torch.library.custom_op("fbct::back_projection", mutates_args=())
def back_projection(projection: torch.Tensor, angle: List[float], parameter: List[float], vol_shape: List[int],
                    proj_shape: List[int]) -> torch.Tensor:
    """The FDK op: nZ = nRows = 1 (FDK/FBP.cc:85,98) makes it the FBP of every slice; the
    slice count comes from the input (the reference ignores parameter[11] and reconstructs
    slice 0 only, FDK/GPURecon_GPU.cu:178-179)."""
    plan = _plan(angle, parameter)
    return _shaped(plan.fdk(_fdk_in(projection, plan, "back_projection")), vol_shape, "back_projection")


# ------------------------------------------------------------------ fake impls

def _fake_to(shape_arg):
    def fake(x, angle, parameter, vol_shape, proj_shape):
        shape = vol_shape if shape_arg == "vol" else proj_shape
        return x.new_empty([int(s) for s in shape], dtype=torch.float32)
    return fake


for _op, _to in ((fp_fan_curve, "proj"), (fp_fan_curve_adjoint, "vol"), (bp_fan_curve, "vol"),
                 (bp_fan_curve_adjoint, "proj"), (fbp_fan_curve, "vol"), (fbp_fan_curve_adjoint, "proj"),
                 (back_projection, "vol")):
    _op.register_fake(_fake_to(_to))


# ------------------------------------------------------------------ autograd

def _setup(ctx, inputs, output):
    x, angle, parameter, vol_shape, proj_shape = inputs
    ctx.attrs = (list(angle), list(parameter), list(vol_shape), list(proj_shape))
    ctx.in_shape = tuple(x.shape)
    ctx.in_dtype = x.dtype


def _backward_with(op):
    def backward(ctx, grad):
        angle, parameter, vol_shape, proj_shape = ctx.attrs
        g = op(grad.contiguous(), angle, parameter, vol_shape, proj_shape)
        return g.reshape(ctx.in_shape).to(ctx.in_dtype), None, None, None, None
    return backward


fp_fan_curve.register_autograd(_backward_with(fp_fan_curve_adjoint), setup_context=_setup)
fp_fan_curve_adjoint.register_autograd(_backward_with(fp_fan_curve), setup_context=_setup)
bp_fan_curve.register_autograd(_backward_with(bp_fan_curve_adjoint), setup_context=_setup)
bp_fan_curve_adjoint.register_autograd(_backward_with(bp_fan_curve), setup_context=_setup)
fbp_fan_curve.register_autograd(_backward_with(fbp_fan_curve_adjoint), setup_context=_setup)
fbp_fan_curve_adjoint.register_autograd(_backward_with(fbp_fan_curve), setup_context=_setup)
back_projection.register_autograd(_backward_with(fbp_fan_curve_adjoint), setup_context=_setup)

OPS = {
    "fp_fan_curve": fp_fan_curve, "fp_fan_curve_adjoint": fp_fan_curve_adjoint,
    "bp_fan_curve": bp_fan_curve, "bp_fan_curve_adjoint": bp_fan_curve_adjoint,
    "fbp_fan_curve": fbp_fan_curve, "fbp_fan_curve_adjoint": fbp_fan_curve_adjoint,
    "back_projection": back_projection,
}
