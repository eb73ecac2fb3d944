"""Host-side plan object: geometry + device execution through the C ABI.

A :class:`FanBeamPlan` is built from the reference op attributes -- the 16-float
``parameter`` vector (layout of ``utils.param_set``, utils.py:56-88) and the view
``angle`` list in radians (``utils.angle``, utils.py:49-54) -- and executes the
operators on CUDA (HIP) tensors on the caller's current stream.  All arithmetic
runs in the gfx950 kernels of ``libfbct.so``; there is no CPU path.
"""
from __future__ import annotations

import collections
import ctypes
import threading

import numpy as np
import torch

from . import _lib

__all__ = ["FanBeamPlan", "get_plan"]


def _as_param_array(parameter) -> np.ndarray:
    if isinstance(parameter, torch.Tensor):
        parameter = parameter.detach().cpu().numpy()
    p = np.ascontiguousarray(np.asarray(parameter, dtype=np.float32).reshape(-1))
    if p.size != 16:
        raise ValueError("parameter must hold 16 values (utils.param_set layout), got %d" % p.size)
    return p


def _as_angle_array(angle) -> np.ndarray:
    if isinstance(angle, torch.Tensor):
        angle = angle.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(angle, dtype=np.float32).reshape(-1))


# fbct_dtype codes of the storage types the operators accept
_DTYPE = {torch.float32: 0, torch.bfloat16: 1}
_STORAGE = (torch.float32, torch.bfloat16)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class FanBeamPlan:
    """Immutable fan-beam geometry with FP64-precomputed tables (one ``fbct_plan``)."""

    def __init__(self, parameter, angle):
        lib = _lib.load()
        self.parameter = _as_param_array(parameter)
        self.angle = _as_angle_array(angle)
        h = ctypes.c_void_p()
        st = lib.fbct_plan_create(_fptr(self.parameter), 16, _fptr(self.angle), int(self.angle.size),
                                  ctypes.byref(h))
        _lib.check(st, "fbct_plan_create")
        self._h = h
        g = _lib.FanGeom()
        _lib.check(lib.fbct_plan_geom(self._h, ctypes.byref(g)), "fbct_plan_geom")
        self.geom = g.as_dict()
        self.n_views = g.nViews
        self.n_cols = g.nCols
        self.nx = g.nX
        self.ny = g.nY
        self.n_samp = g.nSamp

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().fbct_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    # ----------------------------------------------------------- host queries
    @property
    def image_shape(self):
        return (self.ny, self.nx)

    @property
    def sino_shape(self):
        return (self.n_views, self.n_cols)

    @property
    def fp_samples(self) -> int:
        """FP samples per slice (sum over rays of floor(chord/0.125 px))."""
        return int(_lib.load().fbct_plan_fp_samples(self._h))

    def ray_table(self):
        """(N [nViews, nCols] int32, geo [nViews, nCols, 4] float32 = u0, w0, du, dw)."""
        n = np.zeros((self.n_views, self.n_cols), dtype=np.int32)
        geo = np.zeros((self.n_views, self.n_cols, 4), dtype=np.float32)
        _lib.check(_lib.load().fbct_plan_ray_table(
            self._h, n.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _fptr(geo)), "fbct_plan_ray_table")
        return n, geo

    def fp_bundles(self, cfg: int):
        """FP bundle tables of configuration cfg: (info dict, major [nViews, nb] uint8,
        slabs [nViews, nb, 2] int16 = first and last slab with samples, rows
        [nViews, nb, nslab, 2] int16 = staged tile rows [lo, hi])."""
        lib = _lib.load()
        info = np.zeros(5, dtype=np.int32)
        _lib.check(lib.fbct_plan_fp_bundles(self._h, cfg, info.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            None, None, None), "fbct_plan_fp_bundles")
        rays, slab, nslab, nb, cap = (int(v) for v in info)
        major = np.zeros((self.n_views, nb), dtype=np.uint8)
        slabs = np.zeros((self.n_views, nb, 2), dtype=np.int16)
        rows = np.zeros((self.n_views, nb, nslab, 2), dtype=np.int16)
        _lib.check(lib.fbct_plan_fp_bundles(self._h, cfg, None, major.ctypes.data, slabs.ctypes.data,
                                            rows.ctypes.data), "fbct_plan_fp_bundles")
        return dict(rays=rays, slab=slab, nslab=nslab, nb=nb, cap=cap), major, slabs, rows

    def ramp_taps(self) -> np.ndarray:
        h = np.zeros(2 * self.n_cols - 1, dtype=np.float64)
        _lib.check(_lib.load().fbct_plan_ramp_taps(self._h, h.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "fbct_plan_ramp_taps")
        return h

    # ------------------------------------------------------------- execution
    def _prep(self, x: torch.Tensor, shape, what: str, dtypes=(torch.float32,)) -> torch.Tensor:
        if not isinstance(x, torch.Tensor):
            raise TypeError("%s: expected a torch.Tensor" % what)
        if not x.is_cuda:
            raise ValueError("%s: tensor must be on a HIP device (got %s); this library has no CPU path"
                             % (what, x.device))
        if x.dtype not in dtypes:
            raise TypeError("%s: expected %s, got %s" % (what, " or ".join(map(str, dtypes)), x.dtype))
        if x.dim() == 2:
            x = x.unsqueeze(0)
        if x.dim() != 3 or tuple(x.shape[1:]) != tuple(shape):
            raise ValueError("%s: expected [n, %d, %d], got %s" % (what, shape[0], shape[1], tuple(x.shape)))
        return x.contiguous()

    def _run(self, fn_name: str, x: torch.Tensor, out_shape, extra_ws: bool = False) -> torch.Tensor:
        lib = _lib.load()
        n = x.shape[0]
        out = torch.empty((n,) + tuple(out_shape), dtype=torch.float32, device=x.device)
        if n == 0:
            return out
        with torch.cuda.device(x.device):
            stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
            fn = getattr(lib, fn_name)
            if extra_ws:
                ws_bytes = int(lib.fbct_fbp_workspace_size(self._h, n))
                ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=x.device)
                st = fn(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), int(n),
                        ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(ws_bytes), stream)
            else:
                st = fn(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), int(n), stream)
        _lib.check(st, fn_name)
        return out

    def _run_ex(self, fn_name: str, x: torch.Tensor, out_shape, out_dtype) -> torch.Tensor:
        lib = _lib.load()
        n = x.shape[0]
        out = torch.empty((n,) + tuple(out_shape), dtype=out_dtype, device=x.device)
        if n == 0:
            return out
        with torch.cuda.device(x.device):
            stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
            st = getattr(lib, fn_name)(self._h, ctypes.c_void_p(x.data_ptr()), _DTYPE[x.dtype],
                                       ctypes.c_void_p(out.data_ptr()), _DTYPE[out_dtype], int(n), stream)
        _lib.check(st, fn_name)
        return out

    def fp(self, image: torch.Tensor, out_dtype=torch.float32) -> torch.Tensor:
        """image [n, nY, nX] -> sinogram [n, nViews, nCols].

        float32 or bfloat16 storage for either side (arithmetic is FP32)."""
        x = self._prep(image, self.image_shape, "fp", _STORAGE)
        if x.dtype == torch.float32 and out_dtype == torch.float32:
            return self._run("fbct_fp", x, self.sino_shape)
        return self._run_ex("fbct_fp_ex", x, self.sino_shape, out_dtype)

    def fp_residual(self, image: torch.Tensor, target: torch.Tensor, out_dtype=torch.bfloat16) -> torch.Tensor:
        """FP(image) - target in one pass (target: float32 sinograms), stored as out_dtype;
        bitwise equal to ``(self.fp(image) - target).to(out_dtype)``."""
        x = self._prep(image, self.image_shape, "fp_residual", _STORAGE)
        y = self._prep(target, self.sino_shape, "fp_residual target")
        if y.shape[0] != x.shape[0] or y.device != x.device:
            raise ValueError("fp_residual: target must have the image batch's size and device")
        lib = _lib.load()
        n = x.shape[0]
        out = torch.empty((n,) + tuple(self.sino_shape), dtype=out_dtype, device=x.device)
        if n == 0:
            return out
        with torch.cuda.device(x.device):
            stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
            st = lib.fbct_fp_residual(self._h, ctypes.c_void_p(x.data_ptr()), _DTYPE[x.dtype],
                                      ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                      _DTYPE[out_dtype], int(n), stream)
        _lib.check(st, "fbct_fp_residual")
        return out

    def fp_adjoint(self, sino: torch.Tensor, out_dtype=torch.float32) -> torch.Tensor:
        """Exact transpose of :meth:`fp`: sinogram -> image (float32 or bfloat16 storage)."""
        y = self._prep(sino, self.sino_shape, "fp_adjoint", _STORAGE)
        if y.dtype == torch.float32 and out_dtype == torch.float32:
            return self._run("fbct_fp_adjoint", y, self.image_shape)
        return self._run_ex("fbct_fp_adjoint_ex", y, self.image_shape, out_dtype)

    def bp(self, sino: torch.Tensor) -> torch.Tensor:
        return self._run("fbct_bp", self._prep(sino, self.sino_shape, "bp"), self.image_shape)

    def bp_adjoint(self, image: torch.Tensor) -> torch.Tensor:
        return self._run("fbct_bp_adjoint", self._prep(image, self.image_shape, "bp_adjoint"), self.sino_shape)

    def fbp_weighted_bp(self, filtered_sino: torch.Tensor) -> torch.Tensor:
        """The 1/L^2-weighted backprojection step of FBP on an already filtered sinogram."""
        return self._run("fbct_fbp_backproject", self._prep(filtered_sino, self.sino_shape, "fbp_weighted_bp"),
                         self.image_shape)

    def fbp_weighted_bp_adjoint(self, image: torch.Tensor) -> torch.Tensor:
        return self._run("fbct_fbp_backproject_adjoint",
                         self._prep(image, self.image_shape, "fbp_weighted_bp_adjoint"), self.sino_shape)

    def fbp(self, sino: torch.Tensor, cosine: bool = False, short_scan: bool = False) -> torch.Tensor:
        """FBP of sinograms [n, nViews, nCols] -> images [n, nY, nX].  The defaults are the
        reference FBP op; cosine=True adds the equiangular cos(gamma) pre-weight the
        reference omits, short_scan=True Parker's weights (fbct_fbp_ex flags)."""
        x = self._prep(sino, self.sino_shape, "fbp")
        flags = (_lib.FBP_COSINE if cosine else 0) | (_lib.FBP_SHORT_SCAN if short_scan else 0)
        if flags == 0:
            return self._run("fbct_fbp", x, self.image_shape, extra_ws=True)
        lib = _lib.load()
        n = x.shape[0]
        out = torch.empty((n,) + tuple(self.image_shape), dtype=torch.float32, device=x.device)
        if n == 0:
            return out
        with torch.cuda.device(x.device):
            stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
            ws_bytes = int(lib.fbct_fbp_workspace_size(self._h, n))
            ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=x.device)
            st = lib.fbct_fbp_ex(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), int(n),
                                 ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(ws_bytes), ctypes.c_uint(flags),
                                 stream)
        _lib.check(st, "fbct_fbp_ex")
        return out

    def fbp_adjoint(self, image: torch.Tensor) -> torch.Tensor:
        return self._run("fbct_fbp_adjoint", self._prep(image, self.image_shape, "fbp_adjoint"),
                         self.sino_shape, extra_ws=True)

    def parker_weights(self) -> np.ndarray:
        """Short-scan weights [nViews, nCols] (Parker's weights times 2)."""
        w = np.zeros((self.n_views, self.n_cols), dtype=np.float32)
        _lib.check(_lib.load().fbct_plan_parker_weights(self._h, _fptr(w)), "fbct_plan_parker_weights")
        return w

    def fbp_short_scan(self, sino: torch.Tensor) -> torch.Tensor:
        """FBP of a short scan (180 degrees plus the fan angle) with Parker's weights."""
        return self._run("fbct_fbp_short_scan", self._prep(sino, self.sino_shape, "fbp_short_scan"),
                         self.image_shape, extra_ws=True)

    def fdk(self, proj: torch.Tensor) -> torch.Tensor:
        return self._run("fbct_fdk", self._prep(proj, self.sino_shape, "fdk"), self.image_shape, extra_ws=True)

    def ramp_filter(self, sino: torch.Tensor, mode: str = "mfma") -> torch.Tensor:
        """Ramp filter of every detector row.  mode "mfma" (BF16 matrix cores, three-way
        split operands) or "fp64" (direct convolution with FP64 accumulation)."""
        if mode not in ("mfma", "fp64"):
            raise ValueError("ramp_filter: mode must be 'mfma' or 'fp64'")
        x = self._prep(sino, self.sino_shape, "ramp_filter")
        lib = _lib.load()
        out = torch.empty_like(x)
        if x.shape[0] == 0:
            return out
        with torch.cuda.device(x.device):
            stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
            st = lib.fbct_ramp_filter_ex(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                         ctypes.c_int64(x.shape[0] * self.n_views), 1 if mode == "fp64" else 0,
                                         stream)
        _lib.check(st, "fbct_ramp_filter")
        return out


_cache_lock = threading.Lock()
_cache: "collections.OrderedDict[bytes, FanBeamPlan]" = collections.OrderedDict()
_CACHE_SIZE = 32


def get_plan(parameter, angle) -> FanBeamPlan:
    """Plan for (parameter, angle), cached by content (the op attributes are immutable)."""
    p = _as_param_array(parameter)
    a = _as_angle_array(angle)
    key = p.tobytes() + b"|" + a.tobytes()
    with _cache_lock:
        plan = _cache.get(key)
        if plan is not None:
            _cache.move_to_end(key)
            return plan
    plan = FanBeamPlan(p, a)
    with _cache_lock:
        _cache[key] = plan
        while len(_cache) > _CACHE_SIZE:
            _cache.popitem(last=False)
    return plan
