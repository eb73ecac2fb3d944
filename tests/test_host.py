"""CPU tests of the host side of libfbct: the library loads and exports every symbol the
public header declares, validates its arguments without touching a GPU, and builds the
same FP64 geometry tables as the oracle; plus the caller compatibility layer."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from conftest import REPO
from fanbeamctrecon_tensorflowop_amd import _lib, compat, workloads
from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan


def header_functions():
    src = open(os.path.join(REPO, "include", "fbct.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(fbct_[a-z_0-9]+)\s*\(", src, re.M)))


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes signatures cover exactly the header
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == names


def test_version_and_status_strings():
    lib = _lib.load()
    assert lib.fbct_version() == 100
    assert lib.fbct_status_string(0) == b"ok"
    assert lib.fbct_status_string(7).startswith(b"workspace")


def _geom(p, n=16):
    g = _lib.FanGeom()
    st = _lib.load().fbct_geom_from_params(p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, ctypes.byref(g))
    return st, g


def test_geom_decode_matches_oracle(oracle_mod):
    for p, a in (compat.param_set(3), compat.angle(984, 0)), workloads.clinical_scan(), workloads.small_scan():
        st, g = _geom(p)
        assert st == 0
        o = oracle_mod.Oracle(p, a)
        for name, _ in _lib.FanGeom._fields_:
            assert getattr(g, name) == getattr(o.g, name), name


@pytest.mark.parametrize("idx,val,status", [
    (2, 0, 2),        # no views
    (4, -3, 2),       # negative columns
    (7, 0.0, 2),      # zero FOV
    (1, 100.0, 2),    # source inside the FOV (R < rFOV)
    (8, 0.0, 2),      # zero column pitch
])
def test_geom_validation(idx, val, status):
    p = compat.param_set(1).copy()
    p[idx] = val
    st, _ = _geom(p)
    assert st == status
    assert _lib.last_error()


def test_short_parameter_vector_rejected():
    st, _ = _geom(compat.param_set(1), n=12)
    assert st == 1


def test_plan_rejects_angle_count_mismatch():
    p = compat.param_set(1)
    with pytest.raises(_lib.FbctError) as e:
        FanBeamPlan(p, compat.angle(983, 0))
    assert e.value.status == 3


def test_execute_validation_without_gpu():
    lib = _lib.load()
    plan = FanBeamPlan(*workloads.tiny_scan())
    h = plan._h
    null = ctypes.c_void_p(0)
    # zero slices is a no-op that never touches the device
    assert lib.fbct_fp(h, null, null, 0, null) == 0
    assert lib.fbct_bp(h, null, null, 0, null) == 0
    # null data with slices, negative counts, missing workspace
    assert lib.fbct_fp(h, null, null, 1, null) == 1
    assert lib.fbct_fp_adjoint(h, null, null, -1, null) == 1
    assert lib.fbct_fp(None, null, null, 1, null) == 1
    fake = ctypes.c_void_p(0x1000)
    assert lib.fbct_fbp(h, fake, fake, 1, null, 0, null) == 7
    assert lib.fbct_fbp_workspace_size(h, 2) == 2 * plan.n_views * plan.n_cols * 4
    assert lib.fbct_ramp_filter(h, fake, fake, 3, null) == 1  # in == out
    # fused residual: a target is required, storage types are checked
    assert lib.fbct_fp_residual(h, fake, 0, null, fake, 1, 0, null) == 0
    assert lib.fbct_fp_residual(h, fake, 0, null, fake, 1, 2, null) == 1
    assert lib.fbct_fp_residual(h, fake, 0, fake, fake, 5, 2, null) == 1


def test_overlapping_buffers_rejected_without_gpu():
    """Buffers one call reads and writes must be disjoint: the kernels read one while other
    workgroups write the other, so aliasing (a workspace passed as the sinogram, an
    in-place filter) is refused with FBCT_ERR_INVALID_ARGUMENT before any device work."""
    from fanbeamctrecon_tensorflowop_amd.cone import ConePlan
    lib = _lib.load()
    plan = FanBeamPlan(*workloads.tiny_scan())
    h = plan._h
    null = ctypes.c_void_p(0)
    sino_b = plan.n_views * plan.n_cols * 4
    img_b = plan.nx * plan.ny * 4
    a, b, c = 0x10000000, 0x20000000, 0x30000000  # three disjoint fake device ranges
    P = ctypes.c_void_p
    ws = lib.fbct_fbp_workspace_size(h, 1)
    # image and sinogram of one call
    assert lib.fbct_fp(h, P(a), P(a + img_b - 4), 1, null) == 1
    assert lib.fbct_fp_adjoint(h, P(a), P(a), 1, null) == 1
    assert lib.fbct_bp(h, P(a), P(a + 8), 1, null) == 1
    assert lib.fbct_bp_adjoint(h, P(a), P(a), 1, null) == 1
    assert lib.fbct_fp_ex(h, P(a), 1, P(a + img_b // 2 - 2), 1, 1, null) == 1
    # the FBP workspace against the sinogram and the image (forward and transpose)
    assert lib.fbct_fbp(h, P(a), P(b), 1, P(a), ws, null) == 1
    assert lib.fbct_fbp(h, P(a), P(b), 1, P(b + img_b - 4), ws, null) == 1
    assert lib.fbct_fbp_adjoint(h, P(b), P(a), 1, P(a + sino_b - 4), ws, null) == 1
    assert lib.fbct_fbp_adjoint(h, P(b), P(a), 1, P(b), ws, null) == 1
    assert lib.fbct_fbp_ex(h, P(a), P(b), 1, P(a), ws, 1, null) == 1
    assert "must not overlap" in _lib.last_error()
    # the residual against its target
    assert lib.fbct_fp_residual(h, P(a), 0, P(b), P(b), 0, 1, null) == 1
    # cone: in-place row filter, workspace aliasing the projections or the volume
    g, ang = workloads.cone_scan(nx=16, nz=4, dx=1.0, n_views=12, n_cols=24, n_rows=6)
    cp = ConePlan(g, ang)
    proj_b = 12 * 6 * 24 * 4
    vol_b = 4 * 16 * 16 * 4
    cws = lib.fbct_cone_workspace_size(cp._h, 1)
    assert lib.fbct_cone_filter(cp._h, P(a), 0, P(a + proj_b // 2), 1, null) == 1
    assert lib.fbct_cone_backproject(cp._h, P(a), P(a + proj_b - 4), 0, 1, null) == 1
    assert lib.fbct_cone_fdk(cp._h, P(a), 0, P(b), 0, 1, P(a), cws, null) == 1
    assert lib.fbct_cone_fdk(cp._h, P(a), 0, P(b), 0, 1, P(b + vol_b - 4), cws, null) == 1
    assert lib.fbct_cone_fdk(cp._h, P(a), 0, P(a), 0, 1, P(c), cws, null) == 1


def test_uneven_batch_host_side():
    """fbct_batch_* creation, layout, work counts and shape validation (no device calls)."""
    from fanbeamctrecon_tensorflowop_amd.batch import UnevenBatch
    lib = _lib.load()
    scans = workloads.uneven_scans(5, seed=3, sides=(16, 24), views=(10, 20), bins=(12, 30), rfov=8.0, sid=60.0,
                                   sdd=120.0)
    plans = [FanBeamPlan(p, a) for p, a in scans]
    b = UnevenBatch(plans)
    assert lib.fbct_batch_size(b._h) == 5
    assert list(np.diff(b.image_offsets)) == [p.nx * p.ny for p in plans]
    assert list(np.diff(b.sino_offsets)) == [p.n_views * p.n_cols for p in plans]
    assert b.fp_samples == sum(p.fp_samples for p in plans)
    assert b.updates == sum(p.nx * p.ny * p.n_views for p in plans)
    assert b.shape == {"fp_rays": 64, "fp_wgs_per_cu": 0, "adj_tile_x": 16, "adj_wgs_per_cu": 0,
                       "adj_stage_views": 0}
    b.set_shape(fp_rays=32, adj_tile_x=64, fp_wgs_per_cu=1, adj_wgs_per_cu=32, adj_stage_views=8)
    assert b.shape == {"fp_rays": 32, "fp_wgs_per_cu": 1, "adj_tile_x": 64, "adj_wgs_per_cu": 32,
                       "adj_stage_views": 8}
    for bad in ({"fp_rays": 48}, {"adj_tile_x": 12}, {"fp_wgs_per_cu": -1}, {"adj_wgs_per_cu": 33},
                {"adj_stage_views": 2}, {"adj_stage_views": 16}):
        with pytest.raises(_lib.FbctError) as e:
            b.set_shape(**bad)
        assert e.value.status == 1
    assert b.shape["fp_rays"] == 32  # a rejected shape leaves the batch unchanged
    null = ctypes.c_void_p(0)
    assert lib.fbct_batch_fp(b._h, null, 0, null, 0, null) == 1
    assert lib.fbct_batch_fp_adjoint(b._h, ctypes.c_void_p(0x1000), 7, ctypes.c_void_p(0x1000), 0, null) == 1
    out = ctypes.c_void_p()
    assert lib.fbct_batch_create(None, 3, ctypes.byref(out)) == 1
    arr = (ctypes.c_void_p * 2)(plans[0]._h.value, None)
    assert lib.fbct_batch_create(arr, 2, ctypes.byref(out)) == 1
    assert lib.fbct_batch_create(arr, 0, ctypes.byref(out)) == 1


def test_ray_table_matches_oracle(oracle_mod):
    for p, a in (workloads.small_scan(), workloads.clinical_scan()):
        plan = FanBeamPlan(p, a)
        o = oracle_mod.Oracle(p, a)
        n, geo = plan.ray_table()
        assert n.sum() == o.sample_count() == plan.fp_samples
        rng = np.random.default_rng(1)
        for _ in range(300):
            v, c = int(rng.integers(0, len(a))), int(rng.integers(0, plan.n_cols))
            N, vals = o.ray(v, c)
            assert n[v, c] == N
            if N:
                np.testing.assert_array_equal(geo[v, c], np.asarray(vals, dtype=np.float32))


def _fp_bundle_geometries():
    yield "small", workloads.small_scan(), 40
    yield "clinical", workloads.clinical_scan(), 12
    # source close to the FOV (R = 1.3 rFOV), wide fan, off-centre detector, non-square image
    yield "close", workloads.scan(nx=80, ny=64, rfov=40.0, sid=52.0, sdd=110.0, n_cols=200,
                                  dcol_rad=1.6 / 200, n_views=90, dshift=3.5), None


@pytest.mark.parametrize("cfg", [0, 1])
def test_fp_bundle_tables_cover_samples(cfg):
    """Every FP sample, at the FP32 position the kernel computes (fma of k, step, start),
    lies in a slab of its bundle's slab range, and its taps (floor m, floor m + 1) lie in
    that slab's staged rows; no tile exceeds cap rows."""
    for name, (p, a), n_check in _fp_bundle_geometries():
        plan = FanBeamPlan(p, a)
        info, major, slabs, rows = plan.fp_bundles(cfg)
        n, geo = plan.ray_table()
        lo, hi = rows[..., 0].astype(np.int64), rows[..., 1].astype(np.int64)
        used = lo <= hi
        assert (hi - lo + 1)[used].max() <= info["cap"], name
        views = range(plan.n_views) if n_check is None else np.linspace(0, plan.n_views - 1, n_check).astype(int)
        for v in views:
            for b in range(info["nb"]):
                c0, c1 = b * info["rays"], min(plan.n_cols, (b + 1) * info["rays"])
                xm = major[v, b] == 1
                for c in range(c0, c1):
                    N = int(n[v, c])
                    if N == 0:
                        continue
                    u0, w0, du, dw = (np.float32(t) for t in geo[v, c])
                    a0, da, m0, dm = (u0, du, w0, dw) if xm else (w0, dw, u0, du)
                    k = np.arange(N, dtype=np.float64)
                    av = (k * np.float64(da) + np.float64(a0)).astype(np.float32).astype(np.float64)
                    mv = (k * np.float64(dm) + np.float64(m0)).astype(np.float32).astype(np.float64)
                    sl = np.floor((av + 1.0) / info["slab"]).astype(np.int64)
                    assert sl.min() >= slabs[v, b, 0] and sl.max() <= slabs[v, b, 1], (name, v, b, c)
                    fm = np.floor(mv).astype(np.int64)
                    assert (lo[v, b, sl] <= fm).all() and (fm + 1 <= hi[v, b, sl]).all(), (name, v, b, c)


def test_ramp_taps_match_oracle(oracle_mod):
    p, a = workloads.clinical_scan()
    np.testing.assert_allclose(FanBeamPlan(p, a).ramp_taps(), oracle_mod.Oracle(p, a).ramp_taps(), rtol=1e-15,
                               atol=0)


def test_compat_param_set_and_angle():
    p = compat.param_set(2)
    expect = np.array([946.73233346, 538.52, 984, 360, 888, 512, 512, 200, 55.0059 / 888, 0, -55.0059 / 2, 2, 0,
                       13, 14, 15], dtype=np.float32)
    np.testing.assert_array_equal(p, expect)
    a = compat.angle(984, 0)
    np.testing.assert_array_equal(a, (np.arange(984) * (2 * math.pi / 984)).astype(np.float32))
    # the reference adds the start angle (degrees) to radians (utils.py:53)
    np.testing.assert_array_equal(compat.angle(4, 10.0), (10.0 + np.arange(4) * (math.pi / 2)).astype(np.float32))


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libfbct.so"))
    with pytest.raises(_lib.FbctLibraryMissing):
        _lib.load()


def test_tensor_checks_reject_cpu_tensors():
    import torch
    plan = FanBeamPlan(*workloads.tiny_scan())
    with pytest.raises(ValueError, match="no CPU path"):
        plan.fp(torch.zeros(1, 16, 16))


def test_torch_ops_registered_with_fake_shapes():
    """torch.ops.fbct.* exist, propagate the attribute-declared shapes on meta tensors and
    refuse CPU data (no CPU path)."""
    import torch
    from fanbeamctrecon_tensorflowop_amd import ops
    p, a = workloads.tiny_scan(n_slices=2)
    pl, al = p.tolist(), a.tolist()
    vol, proj = [2, 16, 16], [2, 24, 20]
    meta_img = torch.empty(2, 16, 16, device="meta")
    meta_sino = torch.empty(2, 24, 20, device="meta")
    assert torch.ops.fbct.fp_fan_curve(meta_img, al, pl, vol, proj).shape == (2, 24, 20)
    assert torch.ops.fbct.fp_fan_curve_adjoint(meta_sino, al, pl, vol, proj).shape == (2, 16, 16)
    assert torch.ops.fbct.bp_fan_curve(meta_sino, al, pl, vol, proj).shape == (2, 16, 16)
    assert torch.ops.fbct.bp_fan_curve_adjoint(meta_img, al, pl, vol, proj).shape == (2, 24, 20)
    assert torch.ops.fbct.fbp_fan_curve(meta_sino, al, pl, vol, proj).shape == (2, 16, 16)
    assert torch.ops.fbct.fbp_fan_curve_adjoint(meta_img, al, pl, vol, proj).shape == (2, 24, 20)
    assert torch.ops.fbct.back_projection(meta_sino, al, pl, vol, proj).shape == (2, 16, 16)
    with pytest.raises(ValueError, match="no CPU path"):
        ops.fp_fan_curve(torch.zeros(2, 16, 16), a, p, vol, proj)


def test_parker_weights_match_oracle(oracle_mod):
    """fbct_plan_parker_weights (host) equals the oracle's weights to float32 rounding;
    a full scan is accepted too (views past 180 deg + fan get weight 0)."""
    import math
    nv_full = 360
    p, a = workloads.scan(nx=64, rfov=32.0, sid=200.0, sdd=400.0, n_cols=96, dcol_rad=0.4 / 96, n_views=nv_full,
                          dshift=0.75)
    nv = int(math.ceil((math.pi + 0.4 + 0.75 * 2 * 0.4 / 96) / (2 * math.pi / nv_full))) + 2
    ps = p.copy()
    ps[2], ps[3] = nv, nv * 360 / nv_full
    plan = FanBeamPlan(ps, a[:nv])
    w = plan.parker_weights()
    ref = oracle_mod.Oracle(ps, a[:nv]).parker_weights()
    np.testing.assert_allclose(w, ref, rtol=0, atol=3e-7)
    assert w.max() == np.float32(2.0) and w.min() == 0.0
    ps2, _ = workloads.scan(nx=64, rfov=32.0, sid=200.0, sdd=400.0, n_cols=96, dcol_rad=0.4 / 96, n_views=170,
                            angular_range=170.0)
    short = FanBeamPlan(ps2, (np.arange(170) * (math.pi / 180)).astype(np.float32))  # 170 deg of views
    with pytest.raises(_lib.FbctError) as e:
        short.parker_weights()
    assert e.value.status == 2


def test_candidate_atan_polynomial_bound():
    """The FP^T candidate ranges take the fan angle from a degree-7 odd polynomial on
    |q| <= 1/2 and widen the range by kCandAtanErr; check, with the coefficients read from
    the kernel source and evaluated in FP32 as the kernel does (Horner in q^2), that the
    polynomial stays within that bound of atan (an underestimated bound could drop a
    candidate ray and change the adjoint)."""
    import re
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "fanbeamctrecon_tensorflowop_amd", "csrc", "fp_adjoint.hip")).read()
    body = src[src.index("float cand_atan(float q)"):]
    body = body[:body.index("}")]
    coef = [float(v) for v in re.findall(r"([-+]?\d\.\d+e[-+]\d+)f", body)]
    assert len(coef) == 4  # a3, a2, a1, a0 in the order of the Horner nest
    bound = float(re.search(r"kCandAtanErr = ([0-9.e+-]+)f", src).group(1))
    q = np.linspace(-0.5, 0.5, 400001).astype(np.float32)
    s = (q * q).astype(np.float32)
    r = np.float32(coef[0])
    for c in coef[1:]:
        r = (r * s + np.float32(c)).astype(np.float32)
    y = (q * r).astype(np.float32)
    err = np.abs(y.astype(np.float64) - np.arctan(q.astype(np.float64))).max()
    assert err < 0.5 * bound, err


def test_block_candidate_linear_model_bound():
    """The unstaged FP^T tile takes a 4 x 4 cell block's candidate columns from a linear
    model of the fan angle about the block centre and widens the half-width by h^2 / Umin^2
    (radians; h = 1.5 texels, Umin = U - (|cos b| + |sin b|) h the least U over the block).
    The fan angle is the polar angle about the source, whose Hessian has norm 1/r^2 <=
    1/U^2.  Check in FP64, over random views, block centres in the FOV and geometries down
    to a source close to the FOV edge, that the model's error at the block's 16 cell
    centres stays within that bound (a bound too small would drop candidate rays)."""
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(400):
        R = rng.uniform(150.0, 800.0)
        rfov = rng.uniform(0.2, 0.8) * R
        dx = 2.0 * rfov / rng.choice([256, 384, 512, 768, 1024])
        b = rng.uniform(0.0, 2.0 * np.pi)
        c, s = np.cos(b), np.sin(b)
        rad = rfov * np.sqrt(rng.uniform(0.0, 1.0))
        phi = rng.uniform(0.0, 2.0 * np.pi)
        xc, yc = rad * np.cos(phi), rad * np.sin(phi)
        U = R - xc * c - yc * s
        t = xc * s - yc * c
        h = 1.5 * dx
        umin = U - (abs(c) + abs(s)) * h
        if umin <= 0.0:
            continue
        g0 = np.arctan(t / U)
        gx = (U * s + t * c) / (U * U + t * t)
        gy = (t * s - U * c) / (U * U + t * t)
        off = (np.arange(4) - 1.5) * dx
        px, py = np.meshgrid(xc + off, yc + off)
        Up = R - px * c - py * s
        tp = px * s - py * c
        err = np.abs(np.arctan(tp / Up) - (g0 + gx * (px - xc) + gy * (py - yc))).max()
        bound = h * h / (umin * umin)
        worst = max(worst, err / bound)
    assert 0.0 < worst <= 1.0, worst


def test_block_candidate_square_support():
    """The block model's half-width uses the cell's square support: a line of unit direction
    d meets a square cell iff it passes within (|d_x| + |d_y|)/2 texels of its centre,
    taken at the source-to-block-centre direction and widened by 0.7072 per radian of
    direction spread (1.415 h / Rmin over the block, plus the disk window's half-width for
    the ray), capped at the half diagonal; Rmin = max(Umin, sqrt(U^2 + t^2) - 1.415 h)
    bounds the distance to the source over the block.  Check in FP64, without the FP32 margins, that
    every cell's exact fan-angle range (the extremes over its four corners) lies inside the
    window the model gives that cell, over random views, block centres and geometries."""
    rng = np.random.default_rng(11)
    atan_err = 2.0e-6
    worst, narrower = 0.0, []
    for _ in range(600):
        R = rng.uniform(150.0, 800.0)
        rfov = rng.uniform(0.2, 0.8) * R
        dx = 2.0 * rfov / rng.choice([256, 384, 512, 768, 1024])
        b = rng.uniform(0.0, 2.0 * np.pi)
        c, s = np.cos(b), np.sin(b)
        rad = rfov * np.sqrt(rng.uniform(0.0, 1.0))
        phi = rng.uniform(0.0, 2.0 * np.pi)
        xc, yc = rad * np.cos(phi), rad * np.sin(phi)
        U = R - xc * c - yc * s
        t = xc * s - yc * c
        h = 1.5 * dx
        umin = U - (abs(c) + abs(s)) * h
        if umin <= 0.0:
            continue
        rmin = max(umin, np.hypot(U, t) - 1.415 * h)  # least distance to the source

        def half_width(rs):  # radians
            z = rs * dx / rmin
            return z * (1.0 + z * z) + atan_err + h * h / (rmin * rmin) if z < 0.9 else np.inf

        hw_disk = half_width(np.sqrt(0.5))
        ax, ay = abs(t * s - U * c), abs(U * s + t * c)
        rs = min(np.sqrt(0.5), 0.5 * (ax + ay) / np.hypot(U, t) + 0.7072 * (1.415 * h / rmin + hw_disk))
        hw = half_width(rs)
        if not np.isfinite(hw):
            continue
        narrower.append(hw / hw_disk)
        g0 = np.arctan(t / U)
        gx = (U * s + t * c) / (U * U + t * t)
        gy = (t * s - U * c) / (U * U + t * t)
        off = (np.arange(4) - 1.5) * dx
        px, py = np.meshgrid(xc + off, yc + off)
        lin = g0 + gx * (px - xc) + gy * (py - yc)
        for ex in (-0.5, 0.5):
            for ey in (-0.5, 0.5):
                qx, qy = px + ex * dx, py + ey * dx
                ang = np.arctan((qx * s - qy * c) / (R - qx * c - qy * s))
                worst = max(worst, np.abs(ang - lin).max() / hw)
    assert 0.0 < worst <= 1.0, worst
    assert np.mean(narrower) < 0.97, np.mean(narrower)
