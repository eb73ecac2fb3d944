"""Uneven batches: slices of different geometries in one persistent launch.

The reference runs one geometry per op call and loops slices on the host
(FP/transform.cu:233-255).  :class:`UnevenBatch` wraps ``fbct_batch``: slice ``s`` has its
own :class:`FanBeamPlan`, the images and sinograms of all slices live in one packed
device buffer each (offsets from :attr:`UnevenBatch.image_offsets` /
:attr:`UnevenBatch.sino_offsets`), and FP or its matched adjoint over the whole batch is
one kernel launch whose fixed grid strides over the work items of every slice.

:func:`autotune` picks the launch shape (rays per FP workgroup, FP^T pixel tile, grid
size) by timing the candidates on the device, and caches the choice on disk.
:class:`StepGraph` captures a FP -> residual -> FP^T step in a HIP graph.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
from typing import Sequence

import numpy as np
import torch

from . import _lib
from .plan import _DTYPE, _STORAGE, FanBeamPlan

__all__ = ["UnevenBatch", "StepGraph", "autotune", "FP_RAYS", "ADJ_TILE_X", "ADJ_STAGE_VIEWS", "WGS_PER_CU"]

FP_RAYS = (64, 32)
ADJ_TILE_X = (16, 32, 8, 64)
ADJ_STAGE_VIEWS = (0, 4, 8)  # 0: cell tiles read their rays from global memory
WGS_PER_CU = (0, 4, 8, 16, 32)  # 0: one workgroup per work item


class UnevenBatch:
    """FP / FP^T over slices with per-slice geometry (one ``fbct_batch``)."""

    def __init__(self, plans: Sequence[FanBeamPlan]):
        lib = _lib.load()
        self.plans = list(plans)  # referenced: the C batch points at their tables
        n = len(self.plans)
        if n < 1:
            raise ValueError("UnevenBatch needs at least one slice")
        arr = (ctypes.c_void_p * n)(*[p._h.value for p in self.plans])
        h = ctypes.c_void_p()
        _lib.check(lib.fbct_batch_create(arr, n, ctypes.byref(h)), "fbct_batch_create")
        self._h = h
        io = np.zeros(n + 1, dtype=np.int64)
        so = np.zeros(n + 1, dtype=np.int64)
        _lib.check(lib.fbct_batch_layout(h, io.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                         so.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))), "fbct_batch_layout")
        self.image_offsets = io
        self.sino_offsets = so
        fs, up = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(lib.fbct_batch_work(h, ctypes.byref(fs), ctypes.byref(up)), "fbct_batch_work")
        self.fp_samples = int(fs.value)
        self.updates = int(up.value)

    def __len__(self):
        return len(self.plans)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().fbct_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------ launch shape
    @property
    def shape(self) -> dict:
        s = _lib.BatchShape()
        _lib.check(_lib.load().fbct_batch_get_shape(self._h, ctypes.byref(s)), "fbct_batch_get_shape")
        return s.as_dict()

    def set_shape(self, **kw) -> None:
        cur = self.shape
        cur.update(kw)
        s = _lib.BatchShape(**cur)
        _lib.check(_lib.load().fbct_batch_set_shape(self._h, ctypes.byref(s)), "fbct_batch_set_shape")

    def signature(self) -> str:
        """Content hash of the slice geometries (parameter vectors and angle lists)."""
        h = hashlib.sha1()
        for p in self.plans:
            h.update(p.parameter.tobytes())
            h.update(p.angle.tobytes())
        return h.hexdigest()

    # ------------------------------------------------------------ packing
    @property
    def image_numel(self) -> int:
        return int(self.image_offsets[-1])

    @property
    def sino_numel(self) -> int:
        return int(self.sino_offsets[-1])

    def pack_images(self, images, dtype=torch.float32, device="cuda") -> torch.Tensor:
        out = torch.empty(self.image_numel, dtype=dtype, device=device)
        for s, (p, im) in enumerate(zip(self.plans, images)):
            t = torch.as_tensor(im).reshape(p.ny, p.nx)
            out[self.image_offsets[s]:self.image_offsets[s + 1]].copy_(t.reshape(-1))
        return out

    def pack_sinos(self, sinos, dtype=torch.float32, device="cuda") -> torch.Tensor:
        out = torch.empty(self.sino_numel, dtype=dtype, device=device)
        for s, (p, sg) in enumerate(zip(self.plans, sinos)):
            t = torch.as_tensor(sg).reshape(p.n_views, p.n_cols)
            out[self.sino_offsets[s]:self.sino_offsets[s + 1]].copy_(t.reshape(-1))
        return out

    def image(self, packed: torch.Tensor, s: int) -> torch.Tensor:
        p = self.plans[s]
        return packed[self.image_offsets[s]:self.image_offsets[s + 1]].view(p.ny, p.nx)

    def sino(self, packed: torch.Tensor, s: int) -> torch.Tensor:
        p = self.plans[s]
        return packed[self.sino_offsets[s]:self.sino_offsets[s + 1]].view(p.n_views, p.n_cols)

    # ------------------------------------------------------------ execution
    def _check(self, x: torch.Tensor, numel: int, what: str) -> torch.Tensor:
        if not isinstance(x, torch.Tensor) or not x.is_cuda:
            raise ValueError("%s: expected a HIP device tensor; this library has no CPU path" % what)
        if x.dtype not in _STORAGE:
            raise TypeError("%s: storage must be float32 or bfloat16, got %s" % (what, x.dtype))
        if x.numel() != numel:
            raise ValueError("%s: packed buffer has %d elements, the batch needs %d" % (what, x.numel(), numel))
        return x.contiguous()

    def upload(self) -> None:
        """Make the slice plans' and the batch's tables resident on the current device."""
        _lib.check(_lib.load().fbct_batch_upload(self._h), "fbct_batch_upload")

    def fp(self, images: torch.Tensor, out_dtype=torch.float32, out: torch.Tensor = None) -> torch.Tensor:
        """Packed images -> packed sinograms."""
        x = self._check(images, self.image_numel, "batch fp")
        if out is None:
            out = torch.empty(self.sino_numel, dtype=out_dtype, device=x.device)
        self._check(out, self.sino_numel, "batch fp output")
        with torch.cuda.device(x.device):
            st = torch.cuda.current_stream(x.device).cuda_stream
            rc = _lib.load().fbct_batch_fp(self._h, ctypes.c_void_p(x.data_ptr()), _DTYPE[x.dtype],
                                           ctypes.c_void_p(out.data_ptr()), _DTYPE[out.dtype], ctypes.c_void_p(st))
        _lib.check(rc, "fbct_batch_fp")
        return out

    def fp_adjoint(self, sinos: torch.Tensor, out_dtype=torch.float32, out: torch.Tensor = None) -> torch.Tensor:
        """Packed sinograms -> packed images (exact transpose of :meth:`fp`)."""
        y = self._check(sinos, self.sino_numel, "batch fp_adjoint")
        if out is None:
            out = torch.empty(self.image_numel, dtype=out_dtype, device=y.device)
        self._check(out, self.image_numel, "batch fp_adjoint output")
        with torch.cuda.device(y.device):
            st = torch.cuda.current_stream(y.device).cuda_stream
            rc = _lib.load().fbct_batch_fp_adjoint(self._h, ctypes.c_void_p(y.data_ptr()), _DTYPE[y.dtype],
                                                   ctypes.c_void_p(out.data_ptr()), _DTYPE[out.dtype],
                                                   ctypes.c_void_p(st))
        _lib.check(rc, "fbct_batch_fp_adjoint")
        return out


class StepGraph:
    """FP -> residual -> FP^T of an uneven batch captured in one HIP graph.

    ``x`` (packed images) and ``target`` (packed sinograms) are static buffers: write new
    data into them and call :meth:`replay`, which returns the static gradient buffer
    ``FP^T (FP x - target)`` (residual stored as ``storage``)."""

    def __init__(self, batch: UnevenBatch, x: torch.Tensor, target: torch.Tensor, storage=torch.float32):
        self.batch, self.x, self.target, self.storage = batch, x, target, storage
        with torch.cuda.device(x.device):
            batch.upload()
            side = torch.cuda.Stream(device=x.device)
            side.wait_stream(torch.cuda.current_stream(x.device))
            with torch.cuda.stream(side):
                for _ in range(2):
                    self._step()
            torch.cuda.current_stream(x.device).wait_stream(side)
            torch.cuda.synchronize(x.device)
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self.grad = self._step()

    def _step(self) -> torch.Tensor:
        y = self.batch.fp(self.x)
        r = (y - self.target).to(self.storage)
        return self.batch.fp_adjoint(r)

    def replay(self) -> torch.Tensor:
        self.graph.replay()
        return self.grad


# ---------------------------------------------------------------- autotuner

def _cache_path() -> str:
    p = os.environ.get("FBCT_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "fbct_amd", "batch_shapes.json")


def _load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store_cache(path: str, key: str, value: dict) -> None:
    try:
        d = _load_cache(path)
        d[key] = value
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".%d.tmp" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: tuning still applies to this process


def _time(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(batch: UnevenBatch, images: torch.Tensor, sinos: torch.Tensor, reps: int = 2,
             use_cache: bool = True, verbose: bool = False) -> dict:
    """Pick the batch's launch shape by timing every candidate, then set it.

    FP: rays per workgroup in ``FP_RAYS`` x workgroups per CU in ``WGS_PER_CU``;
    FP^T: pixel tile width in ``ADJ_TILE_X`` x workgroups per CU with unstaged cell
    tiles, then, at the best tile width, views per LDS staging batch in
    ``ADJ_STAGE_VIEWS`` x workgroups per CU (the tile width shapes the gather tiles,
    the staging the cell tiles).  The two operators are tuned independently on the given
    packed buffers (their storage types are part of the key).  Results are cached in ``$FBCT_TUNE_CACHE`` (default
    ``~/.cache/fbct_amd/batch_shapes.json``) by device name and geometry signature.
    Returns ``{"shape": ..., "fp_ms": ..., "adj_ms": ..., "cached": bool}``."""
    dev = images.device
    key = "%s|%s|%s|%s" % (torch.cuda.get_device_name(dev), batch.signature(), images.dtype, sinos.dtype)
    path = _cache_path()
    if use_cache:
        hit = _load_cache(path).get(key)
        if hit:
            batch.set_shape(**hit["shape"])
            return dict(hit, cached=True)
    batch.upload()
    fp_out = torch.empty(batch.sino_numel, dtype=torch.float32, device=dev)
    adj_out = torch.empty(batch.image_numel, dtype=torch.float32, device=dev)
    best_fp = None
    for rays in FP_RAYS:
        for w in WGS_PER_CU:
            batch.set_shape(fp_rays=rays, fp_wgs_per_cu=w)
            t = _time(lambda: batch.fp(images, out=fp_out), reps)
            if verbose:
                print("autotune fp rays=%d wgs/cu=%d: %.3f ms" % (rays, w, t), flush=True)
            if best_fp is None or t < best_fp[0]:
                best_fp = (t, rays, w)
    best_adj = None
    for tx in ADJ_TILE_X:
        for w in WGS_PER_CU:
            batch.set_shape(adj_tile_x=tx, adj_wgs_per_cu=w, adj_stage_views=0)
            t = _time(lambda: batch.fp_adjoint(sinos, out=adj_out), reps)
            if verbose:
                print("autotune adj tile=%dx%d wgs/cu=%d: %.3f ms" % (tx, 256 // tx, w, t), flush=True)
            if best_adj is None or t < best_adj[0]:
                best_adj = (t, tx, w, 0)
    for vb in ADJ_STAGE_VIEWS[1:]:
        for w in WGS_PER_CU:
            batch.set_shape(adj_tile_x=best_adj[1], adj_wgs_per_cu=w, adj_stage_views=vb)
            t = _time(lambda: batch.fp_adjoint(sinos, out=adj_out), reps)
            if verbose:
                print("autotune adj stage views=%d wgs/cu=%d: %.3f ms" % (vb, w, t), flush=True)
            if t < best_adj[0]:
                best_adj = (t, best_adj[1], w, vb)
    shape = dict(fp_rays=best_fp[1], fp_wgs_per_cu=best_fp[2], adj_tile_x=best_adj[1], adj_wgs_per_cu=best_adj[2],
                 adj_stage_views=best_adj[3])
    batch.set_shape(**shape)
    result = {"shape": shape, "fp_ms": best_fp[0], "adj_ms": best_adj[0]}
    if use_cache:
        _store_cache(path, key, result)
    return dict(result, cached=False)
