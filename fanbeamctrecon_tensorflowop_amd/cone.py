"""Cone-beam FDK plans (flat or curved detector, nRows rows, nZ-slice volumes).

The reference's cone kernel (FDK/GPURecon_GPU.cu:15-50) is equiangular with a z loop,
but its op pins nZ = nRows = 1 (FDK/FBP.cc:85,98).  :class:`ConePlan` wraps
``fbct_cone_plan``: ``detector="curved"`` is that kernel for any nZ and nRows (the
reference filter, weight 1/(U^2+t^2)); ``detector="flat"`` is the textbook
Feldkamp-Davis-Kress reconstruction for a flat panel (cosine pre-weight, Ram-Lak row
filter, weight 1/U^2).  The row filter runs on the BF16 matrix cores (Toeplitz GEMM with
split operands); projections may be stored in BF16.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .plan import _DTYPE, _STORAGE

__all__ = ["ConePlan", "cone_geom"]

_DETECTOR = {"flat": 0, "curved": 1}


def cone_geom(*, R, D, n_views, d_angle, detector, n_cols, n_rows, d_col, d_row, col_left=None, row_bottom=None,
              d_shift=0.0, nx, ny=None, nz, dx, dy=None, dz, xmin=None, ymin=None, zmin=None, rfov=None,
              short_scan=False) -> dict:
    """Geometry dict (the fields of ``fbct_cone_geom``); centred detector and volume by
    default (colLeft = -nCols dCol / 2, rowBottom = -nRows dRow / 2, corners at -n d / 2,
    rFOV = nX dx / 2).  short_scan: Parker-weighted reconstruction of a scan over 180
    degrees plus the fan angle."""
    ny = nx if ny is None else ny
    dy = dx if dy is None else dy
    g = {
        "R": R, "D": D, "nViews": n_views, "dAngle": d_angle, "detector": _DETECTOR[detector],
        "nCols": n_cols, "nRows": n_rows, "dCol": d_col, "dRow": d_row,
        "colLeft": -0.5 * n_cols * d_col if col_left is None else col_left,
        "rowBottom": -0.5 * n_rows * d_row if row_bottom is None else row_bottom,
        "dShift": d_shift, "nX": nx, "nY": ny, "nZ": nz, "dx": dx, "dy": dy, "dz": dz,
        "xmin": -0.5 * nx * dx if xmin is None else xmin, "ymin": -0.5 * ny * dy if ymin is None else ymin,
        "zmin": -0.5 * nz * dz if zmin is None else zmin, "rFOV": 0.5 * nx * dx if rfov is None else rfov,
        "short_scan": 1 if short_scan else 0,
    }
    # the values the library receives (float32 fields)
    st = _lib.ConeGeom(**g)
    return st.as_dict()


class ConePlan:
    """Immutable cone-beam geometry with FP64-precomputed tables (one ``fbct_cone_plan``)."""

    def __init__(self, geom: dict, angles):
        lib = _lib.load()
        self.geom_struct = _lib.ConeGeom(**geom)
        self.geom = self.geom_struct.as_dict()
        if isinstance(angles, torch.Tensor):
            angles = angles.detach().cpu().numpy()
        self.angles = np.ascontiguousarray(np.asarray(angles, dtype=np.float32).reshape(-1))
        h = ctypes.c_void_p()
        st = lib.fbct_cone_plan_create(ctypes.byref(self.geom_struct),
                                       self.angles.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       int(self.angles.size), ctypes.byref(h))
        _lib.check(st, "fbct_cone_plan_create")
        self._h = h
        g = self.geom
        self.proj_shape = (g["nViews"], g["nRows"], g["nCols"])
        self.vol_shape = (g["nZ"], g["nY"], g["nX"])

    @classmethod
    def from_params(cls, parameter, angles, n_rows: int, n_z: int) -> "ConePlan":
        """The reference FDK op geometry (16-float parameter vector) with n_rows rows and n_z slices."""
        lib = _lib.load()
        p = np.ascontiguousarray(np.asarray(parameter, dtype=np.float32).reshape(-1))
        g = _lib.ConeGeom()
        _lib.check(lib.fbct_cone_geom_from_params(p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(p.size),
                                                  int(n_rows), int(n_z), ctypes.byref(g)),
                   "fbct_cone_geom_from_params")
        return cls(g.as_dict(), angles)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().fbct_cone_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def updates(self) -> int:
        """Voxel-view updates of one volume (nX nY nZ nViews)."""
        g = self.geom
        return g["nX"] * g["nY"] * g["nZ"] * g["nViews"]

    def _prep(self, x, shape, what):
        if not isinstance(x, torch.Tensor) or not x.is_cuda:
            raise ValueError("%s: expected a HIP device tensor; this library has no CPU path" % what)
        if x.dtype not in _STORAGE:
            raise TypeError("%s: storage must be float32 or bfloat16, got %s" % (what, x.dtype))
        if x.dim() == len(shape):
            x = x.unsqueeze(0)
        if tuple(x.shape[1:]) != tuple(shape):
            raise ValueError("%s: expected [n, %s], got %s" % (what, ", ".join(map(str, shape)), tuple(x.shape)))
        return x.contiguous()

    def _stream(self, dev):
        return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)

    def filter(self, proj: torch.Tensor) -> torch.Tensor:
        """Pre-weight + ramp filter of every (view, row) line -> FP32."""
        x = self._prep(proj, self.proj_shape, "cone filter")
        out = torch.empty(x.shape, dtype=torch.float32, device=x.device)
        with torch.cuda.device(x.device):
            st = _lib.load().fbct_cone_filter(self._h, ctypes.c_void_p(x.data_ptr()), _DTYPE[x.dtype],
                                              ctypes.c_void_p(out.data_ptr()), int(x.shape[0]), self._stream(x.device))
        _lib.check(st, "fbct_cone_filter")
        return out

    def backproject(self, filtered: torch.Tensor, out_dtype=torch.float32) -> torch.Tensor:
        """Weighted cone-beam backprojection of filtered projections -> volumes."""
        x = self._prep(filtered, self.proj_shape, "cone backproject")
        if x.dtype != torch.float32:
            raise TypeError("cone backproject: filtered projections are float32")
        out = torch.empty((x.shape[0],) + self.vol_shape, dtype=out_dtype, device=x.device)
        with torch.cuda.device(x.device):
            st = _lib.load().fbct_cone_backproject(self._h, ctypes.c_void_p(x.data_ptr()),
                                                   ctypes.c_void_p(out.data_ptr()), _DTYPE[out_dtype],
                                                   int(x.shape[0]), self._stream(x.device))
        _lib.check(st, "fbct_cone_backproject")
        return out

    def fdk(self, proj: torch.Tensor, out_dtype=torch.float32, workspace: torch.Tensor = None) -> torch.Tensor:
        """FDK of projections [n, nViews, nRows, nCols] (float32 or bfloat16) -> [n, nZ, nY, nX]."""
        x = self._prep(proj, self.proj_shape, "cone fdk")
        n = x.shape[0]
        out = torch.empty((n,) + self.vol_shape, dtype=out_dtype, device=x.device)
        if n == 0:
            return out
        lib = _lib.load()
        ws_bytes = int(lib.fbct_cone_workspace_size(self._h, n))
        if workspace is None or workspace.numel() * workspace.element_size() < ws_bytes:
            workspace = torch.empty(ws_bytes, dtype=torch.uint8, device=x.device)
        with torch.cuda.device(x.device):
            st = lib.fbct_cone_fdk(self._h, ctypes.c_void_p(x.data_ptr()), _DTYPE[x.dtype],
                                   ctypes.c_void_p(out.data_ptr()), _DTYPE[out_dtype], int(n),
                                   ctypes.c_void_p(workspace.data_ptr()), ctypes.c_size_t(ws_bytes),
                                   self._stream(x.device))
        _lib.check(st, "fbct_cone_fdk")
        return out
