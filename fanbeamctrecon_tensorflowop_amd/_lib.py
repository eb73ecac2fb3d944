"""ctypes binding of libfbct.so (the C ABI declared in include/fbct.h).

The library is built in-tree by ``fanbeamctrecon_tensorflowop_amd.build`` and loaded
from this package directory.  There is no fallback: if the library is missing,
importing an operator raises ``FbctLibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os
import threading

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# FBCT_LIBRARY names another build of the same library (the host-sanitizer build of
# tests/run_sanitizers.py); by default the in-tree libfbct.so is loaded.
LIB_PATH = os.environ.get("FBCT_LIBRARY") or os.path.join(_PKG_DIR, "libfbct.so")

_lock = threading.Lock()
_lib = None


class FbctLibraryMissing(RuntimeError):
    pass


class FbctError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__("%s failed: %s (status %d: %s)" % (where, message, status, status_name(status)))
        self.status = status


_STATUS = {
    0: "ok", 1: "invalid argument", 2: "invalid geometry", 3: "shape mismatch",
    4: "unsupported", 5: "out of memory", 6: "device error", 7: "workspace",
}


def status_name(s: int) -> str:
    return _STATUS.get(int(s), "unknown")


class FanGeom(ctypes.Structure):
    """Mirror of ``fbct_fan_geom``."""
    _fields_ = [
        ("D", ctypes.c_float), ("R", ctypes.c_float), ("nViews", ctypes.c_int32),
        ("dRange", ctypes.c_float), ("dAngle", ctypes.c_float), ("nCols", ctypes.c_int32),
        ("nX", ctypes.c_int32), ("nY", ctypes.c_int32), ("rFOV", ctypes.c_float),
        ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("xmin", ctypes.c_float),
        ("ymin", ctypes.c_float), ("dCol", ctypes.c_float), ("dShift", ctypes.c_float),
        ("dLeft", ctypes.c_float), ("nSamp", ctypes.c_int32), ("startAngle", ctypes.c_float),
        ("detBottom", ctypes.c_float), ("dRow", ctypes.c_float), ("dz", ctypes.c_float),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchShape(ctypes.Structure):
    """Mirror of ``fbct_batch_shape``."""
    _fields_ = [("fp_rays", ctypes.c_int32), ("fp_wgs_per_cu", ctypes.c_int32),
                ("adj_tile_x", ctypes.c_int32), ("adj_wgs_per_cu", ctypes.c_int32),
                ("adj_stage_views", ctypes.c_int32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class ConeGeom(ctypes.Structure):
    """Mirror of ``fbct_cone_geom``."""
    _fields_ = [
        ("R", ctypes.c_float), ("D", ctypes.c_float), ("nViews", ctypes.c_int32), ("dAngle", ctypes.c_float),
        ("detector", ctypes.c_int32), ("nCols", ctypes.c_int32), ("nRows", ctypes.c_int32),
        ("dCol", ctypes.c_float), ("dRow", ctypes.c_float), ("colLeft", ctypes.c_float),
        ("rowBottom", ctypes.c_float), ("dShift", ctypes.c_float), ("nX", ctypes.c_int32),
        ("nY", ctypes.c_int32), ("nZ", ctypes.c_int32), ("dx", ctypes.c_float), ("dy", ctypes.c_float),
        ("dz", ctypes.c_float), ("xmin", ctypes.c_float), ("ymin", ctypes.c_float), ("zmin", ctypes.c_float),
        ("rFOV", ctypes.c_float), ("short_scan", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) of every symbol the header declares.
FBP_COSINE, FBP_SHORT_SCAN = 1, 2  # fbct_fbp_ex flags

_P = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.c_int
SIGNATURES = [
    ("fbct_version", _I, []),
    ("fbct_status_string", ctypes.c_char_p, [_I]),
    ("fbct_last_error", ctypes.c_char_p, []),
    ("fbct_geom_from_params", _I, [_F, _I, ctypes.POINTER(FanGeom)]),
    ("fbct_plan_create", _I, [_F, _I, _F, _I, ctypes.POINTER(_P)]),
    ("fbct_plan_destroy", _I, [_P]),
    ("fbct_plan_geom", _I, [_P, ctypes.POINTER(FanGeom)]),
    ("fbct_plan_fp_samples", ctypes.c_int64, [_P]),
    ("fbct_plan_ray_table", _I, [_P, ctypes.POINTER(ctypes.c_int32), _F]),
    ("fbct_plan_fp_bundles", _I, [_P, _I, ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p]),
    ("fbct_plan_ramp_taps", _I, [_P, ctypes.POINTER(ctypes.c_double)]),
    ("fbct_plan_upload", _I, [_P]),
    ("fbct_plan_parker_weights", _I, [_P, _F]),
    ("fbct_fbp_short_scan", _I, [_P, _P, _P, _I, _P, ctypes.c_size_t, _P]),
    ("fbct_fp", _I, [_P, _P, _P, _I, _P]),
    ("fbct_fp_adjoint", _I, [_P, _P, _P, _I, _P]),
    ("fbct_fp_ex", _I, [_P, _P, _I, _P, _I, _I, _P]),
    ("fbct_fp_residual", _I, [_P, _P, _I, _P, _P, _I, _I, _P]),
    ("fbct_fp_adjoint_ex", _I, [_P, _P, _I, _P, _I, _I, _P]),
    ("fbct_bp", _I, [_P, _P, _P, _I, _P]),
    ("fbct_bp_adjoint", _I, [_P, _P, _P, _I, _P]),
    ("fbct_fbp_backproject", _I, [_P, _P, _P, _I, _P]),
    ("fbct_fbp_backproject_adjoint", _I, [_P, _P, _P, _I, _P]),
    ("fbct_ramp_filter", _I, [_P, _P, _P, ctypes.c_int64, _P]),
    ("fbct_ramp_filter_ex", _I, [_P, _P, _P, ctypes.c_int64, _I, _P]),
    ("fbct_fbp_workspace_size", ctypes.c_size_t, [_P, _I]),
    ("fbct_fbp", _I, [_P, _P, _P, _I, _P, ctypes.c_size_t, _P]),
    ("fbct_fbp_ex", _I, [_P, _P, _P, _I, _P, ctypes.c_size_t, ctypes.c_uint, _P]),
    ("fbct_fbp_noalloc", _I, [_P, _P, _P, _I, _P]),
    ("fbct_fbp_adjoint", _I, [_P, _P, _P, _I, _P, ctypes.c_size_t, _P]),
    ("fbct_fdk", _I, [_P, _P, _P, _I, _P, ctypes.c_size_t, _P]),
    ("fbct_batch_create", _I, [ctypes.POINTER(_P), _I, ctypes.POINTER(_P)]),
    ("fbct_batch_destroy", _I, [_P]),
    ("fbct_batch_size", _I, [_P]),
    ("fbct_batch_layout", _I, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    ("fbct_batch_work", _I, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    ("fbct_batch_get_shape", _I, [_P, ctypes.POINTER(BatchShape)]),
    ("fbct_batch_set_shape", _I, [_P, ctypes.POINTER(BatchShape)]),
    ("fbct_batch_upload", _I, [_P]),
    ("fbct_batch_fp", _I, [_P, _P, _I, _P, _I, _P]),
    ("fbct_batch_fp_adjoint", _I, [_P, _P, _I, _P, _I, _P]),
    ("fbct_cone_geom_from_params", _I, [_F, _I, _I, _I, ctypes.POINTER(ConeGeom)]),
    ("fbct_cone_plan_create", _I, [ctypes.POINTER(ConeGeom), _F, _I, ctypes.POINTER(_P)]),
    ("fbct_cone_plan_destroy", _I, [_P]),
    ("fbct_cone_workspace_size", ctypes.c_size_t, [_P, _I]),
    ("fbct_cone_filter", _I, [_P, _P, _I, _P, _I, _P]),
    ("fbct_cone_backproject", _I, [_P, _P, _P, _I, _I, _P]),
    ("fbct_cone_fdk", _I, [_P, _P, _I, _P, _I, _I, _P, ctypes.c_size_t, _P]),
]


def load():
    """Load libfbct.so (importing torch first so one HIP runtime serves both)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise FbctLibraryMissing(
                "libfbct.so not found at %s; build it with "
                "`python -m fanbeamctrecon_tensorflowop_amd.build`" % LIB_PATH)
        try:
            import torch  # noqa: F401  (its libamdhip64 is then the one the library binds to)
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def last_error() -> str:
    msg = load().fbct_last_error()
    return msg.decode() if msg else ""


def check(status: int, where: str) -> None:
    if status != 0:
        raise FbctError(int(status), where, last_error())
