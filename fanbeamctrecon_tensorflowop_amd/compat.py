"""Caller compatibility layer: counterparts of the reference's ``utils.py`` and the
fields of ``config.py`` that ``utils.param_set`` consumes.

``angle`` and ``param_set`` return the float32 attribute arrays the reference wraps
in TensorProtos (utils.py:49-88); ``ForwardProjection``, ``BackProjection`` and
``FilteredBackProjection`` call the drop-in operators with the same keywords
(utils.py:15-46).  Quirks are kept on purpose, because callers depend on the
values: ``angle`` adds the start angle (given in degrees) to radians (utils.py:53),
``param_set`` folds the detector shift into dLeft (utils.py:63) which BP/FBP then
apply once more (BP/transform.cu:85), and entries 13-15 keep np.arange's values.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np

from . import ops

__all__ = ["config", "default_config", "angle", "param_set", "make_parameter",
           "ForwardProjection", "BackProjection", "FilteredBackProjection"]


def default_config() -> SimpleNamespace:
    """The scan/image settings of the reference's config.py:13-36 (values as shipped)."""
    return SimpleNamespace(
        ProjectionFilename="./test_prj.raw",
        ReconstructionFilename="./test_img.raw",
        SourceType=1,
        DetectorType=1,
        SourceDetectorDistance=946.73233346,
        SourceAxisDistance=538.52,
        StartAngle=0,
        ProjectionsPerRotation=984,
        DetectorColumnNumber=888,
        DetectorCoverage=200.1,
        DetectorShift=0,
        DetectorRowNumber=1,
        FanAngle=55.0059,
        FullAngle=360,
        ParkWeighting=0,
        NX=512,
        NY=512,
        NZ=1,
        FOVRadius=200,
        DeltaZ=0.625,
        Algorithm="FBP",
        HelicalReconMode=1,
        FilterType=1,
    )


config = default_config()


def angle(nview: int, fstartangle) -> np.ndarray:
    """utils.angle (utils.py:49-54): start + i*2*pi/nview, stored as float32."""
    d_angle = 2.0 * math.pi / nview
    a = np.arange(nview, dtype=np.float32)
    for i in range(nview):
        a[i] = fstartangle + i * d_angle
    return a


def param_set(nimg: int, cfg: SimpleNamespace = None) -> np.ndarray:
    """utils.param_set (utils.py:56-88): the 16-float parameter attribute."""
    cfg = cfg or config
    sdd = cfg.SourceDetectorDistance
    sod = cfg.SourceAxisDistance
    nfan = cfg.FanAngle
    ncol = cfg.DetectorColumnNumber
    dcol = nfan / ncol
    dshift = cfg.DetectorShift
    dleft = -nfan / 2 - dshift * dcol
    p = np.arange(16, dtype=np.float32)
    p[0] = sdd
    p[1] = sod
    p[2] = cfg.ProjectionsPerRotation
    p[3] = cfg.FullAngle
    p[4] = ncol
    p[5] = cfg.NX
    p[6] = cfg.NY
    p[7] = cfg.FOVRadius
    p[8] = dcol
    p[9] = dshift
    p[10] = dleft
    p[11] = 1 * nimg  # Batch_size * nimg
    p[12] = cfg.StartAngle
    return p


def make_parameter(*, sdd, sod, n_views, n_cols, nx, ny, rfov, dcol_deg, dleft_deg=None, dshift=0.0,
                   angular_range=360.0, n_slices=1, start_angle_deg=0.0, det_bottom=13.0, drow=14.0,
                   dz=15.0) -> np.ndarray:
    """Explicit constructor of the same 16-float layout (FDK/FDK_test_training.py:61-77
    fills entries 13-15 with detector bottom, row pitch and slice thickness)."""
    if dleft_deg is None:
        dleft_deg = -0.5 * n_cols * dcol_deg
    p = np.zeros(16, dtype=np.float32)
    p[:] = [sdd, sod, n_views, angular_range, n_cols, nx, ny, rfov, dcol_deg, dshift, dleft_deg,
            n_slices, start_angle_deg, det_bottom, drow, dz]
    return p


def _shapes(nimg, cfg):
    cfg = cfg or config
    nview = cfg.ProjectionsPerRotation
    ncol = cfg.DetectorColumnNumber
    return cfg, nview, ncol, cfg.NX, cfg.NY


def ForwardProjection(img, nimg=1, cfg=None):  # noqa: N802 (reference name)
    """utils.ForwardProjection (utils.py:15-26)."""
    cfg, nview, ncol, nx, ny = _shapes(nimg, cfg)
    return ops.fp_fan_curve(volume=img, angle=angle(nview, cfg.StartAngle), parameter=param_set(nimg, cfg),
                            vol_shape=[nimg, nx, ny], proj_shape=[nimg, nview, ncol])


def BackProjection(sino, nimg=1, cfg=None):  # noqa: N802
    """utils.BackProjection (utils.py:28-36)."""
    cfg, nview, ncol, nx, ny = _shapes(nimg, cfg)
    return ops.bp_fan_curve(projection=sino, angle=angle(nview, cfg.StartAngle), parameter=param_set(nimg, cfg),
                            vol_shape=[nimg, nx, ny], proj_shape=[nimg, nview, ncol])


def FilteredBackProjection(sino, nimg=1, cfg=None):  # noqa: N802
    """utils.FilteredBackProjection (utils.py:38-46)."""
    cfg, nview, ncol, nx, ny = _shapes(nimg, cfg)
    return ops.fbp_fan_curve(projection=sino, angle=angle(nview, cfg.StartAngle), parameter=param_set(nimg, cfg),
                             vol_shape=[nimg, nx, ny], proj_shape=[nimg, nview, ncol])
