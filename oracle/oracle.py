"""ctypes wrapper of the FP64 CPU oracle (oracle/fbct_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker used by tests/, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``.  The product package never imports
this module.

Every function takes numpy arrays in the reference layouts
(image ``[nS, nY, nX]``, sinogram ``[nS, nViews, nCols]``), widens them to
float64 and returns float64 results.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FBCT_ORACLE_LIBRARY names another build of the oracle (tests/run_sanitizers.py).
_LIB_PATH = os.environ.get("FBCT_ORACLE_LIBRARY") or os.path.join(_HERE, "_build", "libfbct_oracle.so")
_lib = None


class OrcGeom(ctypes.Structure):
    _fields_ = [
        ("D", ctypes.c_float), ("R", ctypes.c_float), ("nViews", ctypes.c_int),
        ("dRange", ctypes.c_float), ("dAngle", ctypes.c_float), ("nCols", ctypes.c_int),
        ("nX", ctypes.c_int), ("nY", ctypes.c_int), ("rFOV", ctypes.c_float),
        ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("xmin", ctypes.c_float),
        ("ymin", ctypes.c_float), ("dCol", ctypes.c_float), ("dShift", ctypes.c_float),
        ("dLeft", ctypes.c_float), ("nSamp", ctypes.c_int), ("startAngle", ctypes.c_float),
        ("detBottom", ctypes.c_float), ("dRow", ctypes.c_float), ("dz", ctypes.c_float),
    ]


class OrcCone(ctypes.Structure):
    _fields_ = [
        ("R", ctypes.c_double), ("D", ctypes.c_double), ("nViews", ctypes.c_int), ("dAngle", ctypes.c_double),
        ("detector", ctypes.c_int), ("nCols", ctypes.c_int), ("nRows", ctypes.c_int), ("dCol", ctypes.c_double),
        ("dRow", ctypes.c_double), ("colLeft", ctypes.c_double), ("rowBottom", ctypes.c_double),
        ("dShift", ctypes.c_double), ("nX", ctypes.c_int), ("nY", ctypes.c_int), ("nZ", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("xmin", ctypes.c_double),
        ("ymin", ctypes.c_double), ("zmin", ctypes.c_double), ("rFOV", ctypes.c_double),
        ("short_scan", ctypes.c_int),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if os.environ.get("FBCT_ORACLE_LIBRARY"):
        return _LIB_PATH
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.POINTER
    d = P(ctypes.c_double)
    f = P(ctypes.c_float)
    g = P(OrcGeom)
    i = ctypes.c_int
    lib.orc_decode_params.argtypes = [f, g]
    lib.orc_decode_params.restype = i
    lib.orc_ray_setup.argtypes = [g, ctypes.c_double, i, P(i), d, d, d, d]
    lib.orc_ray_setup.restype = None
    lib.orc_fp_sample_count.argtypes = [g, f]
    lib.orc_fp_sample_count.restype = ctypes.c_longlong
    for name in ("orc_fp", "orc_fp_adjoint", "orc_fbp", "orc_fbp_adjoint"):
        fn = getattr(lib, name)
        fn.argtypes = [g, f, d, d, i, i]
        fn.restype = None
    for name in ("orc_bp", "orc_bp_adjoint"):
        fn = getattr(lib, name)
        fn.argtypes = [g, f, d, d, i, i, i]
        fn.restype = None
    lib.orc_ramp_taps.argtypes = [g, d]
    lib.orc_ramp_taps.restype = None
    lib.orc_ramp_filter.argtypes = [g, d, d, ctypes.c_longlong, i]
    lib.orc_ramp_filter.restype = None
    lib.orc_fdk.argtypes = [g, f, d, d, i, i, i, ctypes.c_double, i]
    lib.orc_fdk.restype = None
    c = P(OrcCone)
    lib.orc_cone_taps.argtypes = [c, d]
    lib.orc_cone_taps.restype = None
    lib.orc_cone_preweight.argtypes = [c, d]
    lib.orc_cone_preweight.restype = None
    lib.orc_cone_filter.argtypes = [c, f, d, d, i, i]
    lib.orc_cone_filter.restype = i
    lib.orc_cone_bp.argtypes = [c, f, d, d, i, i]
    lib.orc_cone_bp.restype = None
    lib.orc_cone_fdk.argtypes = [c, f, d, d, i, i]
    lib.orc_cone_fdk.restype = i
    lib.orc_parker_weights.argtypes = [f, i, d, i, ctypes.c_double, d]
    lib.orc_parker_weights.restype = i
    lib.orc_parker_fan.argtypes = [g, f, d]
    lib.orc_parker_fan.restype = i
    lib.orc_fbp_short_scan.argtypes = [g, f, d, d, i, i]
    lib.orc_fbp_short_scan.restype = i
    lib.orc_fbp_weighted.argtypes = [g, f, d, d, i, i, i, i]
    lib.orc_fbp_weighted.restype = i
    lib.orc_default_threads.argtypes = []
    lib.orc_default_threads.restype = i
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def default_threads() -> int:
    return int(_load().orc_default_threads())


class Oracle:
    """FP64 oracle bound to one geometry (16-float parameter vector + view angles)."""

    def __init__(self, parameter, angles):
        lib = _load()
        self.parameter = np.ascontiguousarray(np.asarray(parameter, dtype=np.float32).reshape(16))
        self.angles = np.ascontiguousarray(np.asarray(angles, dtype=np.float32).reshape(-1))
        self.g = OrcGeom()
        if lib.orc_decode_params(_fptr(self.parameter), ctypes.byref(self.g)) != 0:
            raise ValueError("invalid geometry parameters")
        if self.angles.size != self.g.nViews:
            raise ValueError("angle list length %d != nViews %d" % (self.angles.size, self.g.nViews))

    # ---------------------------------------------------------------- helpers
    @property
    def img_shape(self):
        return (self.g.nY, self.g.nX)

    @property
    def sino_shape(self):
        return (self.g.nViews, self.g.nCols)

    def _img(self, x):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        if x.ndim == 2:
            x = x[None]
        assert x.shape[1:] == self.img_shape, (x.shape, self.img_shape)
        return x

    def _sino(self, p):
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float64))
        if p.ndim == 2:
            p = p[None]
        assert p.shape[1:] == self.sino_shape, (p.shape, self.sino_shape)
        return p

    # ------------------------------------------------------------- geometry
    def ray(self, view: int, col: int):
        n = ctypes.c_int()
        vals = [ctypes.c_double() for _ in range(4)]
        _load().orc_ray_setup(ctypes.byref(self.g), float(self.angles[view]), int(col),
                              ctypes.byref(n), *[ctypes.byref(v) for v in vals])
        return n.value, tuple(v.value for v in vals)

    def sample_count(self) -> int:
        return int(_load().orc_fp_sample_count(ctypes.byref(self.g), _fptr(self.angles)))

    def ramp_taps(self):
        h = np.zeros(2 * self.g.nCols - 1, dtype=np.float64)
        _load().orc_ramp_taps(ctypes.byref(self.g), _dptr(h))
        return h

    # ------------------------------------------------------------ operators
    def fp(self, img, nthreads=0):
        x = self._img(img)
        out = np.zeros((x.shape[0],) + self.sino_shape, dtype=np.float64)
        _load().orc_fp(ctypes.byref(self.g), _fptr(self.angles), _dptr(x), _dptr(out), x.shape[0], nthreads)
        return out

    def fp_adjoint(self, sino, nthreads=0):
        p = self._sino(sino)
        out = np.zeros((p.shape[0],) + self.img_shape, dtype=np.float64)
        _load().orc_fp_adjoint(ctypes.byref(self.g), _fptr(self.angles), _dptr(p), _dptr(out), p.shape[0], nthreads)
        return out

    def bp(self, sino, weighted=False, nthreads=0):
        p = self._sino(sino)
        out = np.zeros((p.shape[0],) + self.img_shape, dtype=np.float64)
        _load().orc_bp(ctypes.byref(self.g), _fptr(self.angles), _dptr(p), _dptr(out), p.shape[0],
                       1 if weighted else 0, nthreads)
        return out

    def bp_adjoint(self, img, weighted=False, nthreads=0):
        x = self._img(img)
        out = np.zeros((x.shape[0],) + self.sino_shape, dtype=np.float64)
        _load().orc_bp_adjoint(ctypes.byref(self.g), _fptr(self.angles), _dptr(x), _dptr(out), x.shape[0],
                               1 if weighted else 0, nthreads)
        return out

    def ramp_filter(self, sino, nthreads=0):
        p = np.ascontiguousarray(np.asarray(sino, dtype=np.float64))
        assert p.shape[-1] == self.g.nCols
        out = np.zeros_like(p)
        rows = p.size // self.g.nCols
        _load().orc_ramp_filter(ctypes.byref(self.g), _dptr(p), _dptr(out), rows, nthreads)
        return out

    def fbp(self, sino, nthreads=0):
        p = self._sino(sino)
        out = np.zeros((p.shape[0],) + self.img_shape, dtype=np.float64)
        _load().orc_fbp(ctypes.byref(self.g), _fptr(self.angles), _dptr(p), _dptr(out), p.shape[0], nthreads)
        return out

    def fbp_adjoint(self, img, nthreads=0):
        x = self._img(img)
        out = np.zeros((x.shape[0],) + self.sino_shape, dtype=np.float64)
        _load().orc_fbp_adjoint(ctypes.byref(self.g), _fptr(self.angles), _dptr(x), _dptr(out), x.shape[0], nthreads)
        return out

    def parker_weights(self):
        w = np.zeros(self.sino_shape)
        if _load().orc_parker_fan(ctypes.byref(self.g), _fptr(self.angles), _dptr(w)) != 0:
            raise ValueError("short scan: the views do not cover pi + the fan angle")
        return w

    def fbp_short_scan(self, sino, nthreads=0):
        p = self._sino(sino)
        out = np.zeros((p.shape[0],) + self.img_shape, dtype=np.float64)
        if _load().orc_fbp_short_scan(ctypes.byref(self.g), _fptr(self.angles), _dptr(p), _dptr(out), p.shape[0],
                                      nthreads) != 0:
            raise ValueError("short scan: the views do not cover pi + the fan angle")
        return out

    def fbp_weighted(self, sino, cosine=False, short_scan=False, nthreads=0):
        """FBP with the cosine pre-weight and/or Parker's weights (orc_fbp_weighted)."""
        p = self._sino(sino)
        out = np.zeros((p.shape[0],) + self.img_shape, dtype=np.float64)
        if _load().orc_fbp_weighted(ctypes.byref(self.g), _fptr(self.angles), _dptr(p), _dptr(out), p.shape[0],
                                    int(bool(cosine)), int(bool(short_scan)), nthreads) != 0:
            raise ValueError("short scan: the views do not cover pi + the fan angle")
        return out

    def fdk(self, proj, n_rows=1, n_z=1, zmin=None, nthreads=0):
        """proj [nS, nViews, nRows, nCols] (or [nS, nViews, nCols] when n_rows == 1)."""
        p = np.ascontiguousarray(np.asarray(proj, dtype=np.float64))
        nS = p.shape[0]
        p = p.reshape(nS, self.g.nViews, n_rows, self.g.nCols)
        if zmin is None:
            zmin = float(np.float32(-0.5) * np.float32(self.g.dz) * np.float32(n_z))
        out = np.zeros((nS, n_z) + self.img_shape, dtype=np.float64)
        _load().orc_fdk(ctypes.byref(self.g), _fptr(self.angles), _dptr(p), _dptr(out), nS, n_rows, n_z,
                        float(zmin), nthreads)
        return out


class ConeOracle:
    """FP64 cone-beam FDK oracle for one geometry (a dict with the fields of
    ``fbct_cone_geom``; values are taken as float32, as the GPU plan receives them)."""

    _INT = ("nViews", "detector", "nCols", "nRows", "nX", "nY", "nZ", "short_scan")

    def __init__(self, geom: dict, angles):
        _load()
        vals = {k: (int(v) if k in self._INT else float(np.float32(v))) for k, v in geom.items()}
        self.g = OrcCone(**vals)
        self.angles = np.ascontiguousarray(np.asarray(angles, dtype=np.float32).reshape(-1))
        if self.angles.size != self.g.nViews:
            raise ValueError("angle list length %d != nViews %d" % (self.angles.size, self.g.nViews))

    @property
    def proj_shape(self):
        return (self.g.nViews, self.g.nRows, self.g.nCols)

    @property
    def vol_shape(self):
        return (self.g.nZ, self.g.nY, self.g.nX)

    def _proj(self, p):
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float64))
        if p.ndim == 3:
            p = p[None]
        assert p.shape[1:] == self.proj_shape, (p.shape, self.proj_shape)
        return p

    def taps(self):
        h = np.zeros(2 * self.g.nCols - 1)
        _load().orc_cone_taps(ctypes.byref(self.g), _dptr(h))
        return h

    def preweight(self):
        w = np.zeros((self.g.nRows, self.g.nCols))
        _load().orc_cone_preweight(ctypes.byref(self.g), _dptr(w))
        return w

    def filter(self, proj, nthreads=0):
        p = self._proj(proj)
        out = np.zeros_like(p)
        if _load().orc_cone_filter(ctypes.byref(self.g), _fptr(self.angles), _dptr(p), _dptr(out), p.shape[0],
                                   nthreads) != 0:
            raise ValueError("short scan: the views do not cover pi + the fan angle")
        return out

    def backproject(self, q, nthreads=0):
        q = self._proj(q)
        out = np.zeros((q.shape[0],) + self.vol_shape)
        _load().orc_cone_bp(ctypes.byref(self.g), _fptr(self.angles), _dptr(q), _dptr(out), q.shape[0], nthreads)
        return out

    def fdk(self, proj, nthreads=0):
        p = self._proj(proj)
        out = np.zeros((p.shape[0],) + self.vol_shape)
        if _load().orc_cone_fdk(ctypes.byref(self.g), _fptr(self.angles), _dptr(p), _dptr(out), p.shape[0],
                                nthreads) != 0:
            raise ValueError("short scan: the views do not cover pi + the fan angle")
        return out
