"""GPU parity: the gfx950 kernels (called through the C ABI) against the FP64 oracle.

Tolerances (relative L2 against the FP64 oracle on the same float32 inputs):
  FP, BP, FBP, FP^T, BP^T, FBP^T .......... 1e-5   (FP32 arithmetic, FP64 geometry)
  ramp filter ............................. 1e-6   (FP64 accumulation, FP32 output)
  adjoint dot tests <Ax,y> vs <x,A^T y> .... 1e-5 relative
Differences from the reference's texture path: exact FP32 interpolation weights instead
of the texture unit's 8-bit fixed-point weights.
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from fanbeamctrecon_tensorflowop_amd import compat, ops, workloads
from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def small(oracle_mod):
    p, a = workloads.small_scan()
    plan = FanBeamPlan(p, a)
    o = oracle_mod.Oracle(p, a)
    img = workloads.ellipse_phantom(256, 256, 128.0, 8, seed=0)[None]
    return plan, o, img


@pytest.fixture(scope="module")
def clinical(oracle_mod):
    p, a = workloads.clinical_scan()
    plan = FanBeamPlan(p, a)
    o = oracle_mod.Oracle(p, a)
    img = workloads.clinical_phantom(1)
    return plan, o, img


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


# ----------------------------------------------------------------- forward ops

@pytest.mark.parametrize("which", ["small", "clinical"])
def test_fp_matches_oracle(which, request):
    plan, o, img = request.getfixturevalue(which)
    _say("fp %s" % which)
    got = host(plan.fp(dev(img)))
    ref = o.fp(img)
    err = rel_l2(got, ref)
    _say("fp %s rel_l2=%.3e" % (which, err))
    assert err <= TOL


@pytest.mark.parametrize("which", ["small", "clinical"])
def test_bp_matches_oracle(which, request):
    plan, o, img = request.getfixturevalue(which)
    sino = o.fp(img).astype(np.float32)
    _say("bp %s" % which)
    got = host(plan.bp(dev(sino)))
    err = rel_l2(got, o.bp(sino))
    _say("bp %s rel_l2=%.3e" % (which, err))
    assert err <= TOL


@pytest.mark.parametrize("which", ["small", "clinical"])
def test_fbp_matches_oracle(which, request):
    plan, o, img = request.getfixturevalue(which)
    sino = o.fp(img).astype(np.float32)
    _say("fbp %s" % which)
    got = host(plan.fbp(dev(sino)))
    err = rel_l2(got, o.fbp(sino))
    # split the error: the ramp filter alone, and the weighted BP of the oracle-filtered rows
    q_ref = o.ramp_filter(sino)
    e_filt = rel_l2(host(plan.ramp_filter(dev(sino))), q_ref)
    e_wbp = rel_l2(host(plan.fbp_weighted_bp(dev(q_ref))), o.bp(q_ref.astype(np.float32), weighted=True))
    _say("fbp %s rel_l2=%.3e (filter %.3e, weighted bp %.3e)" % (which, err, e_filt, e_wbp))
    assert err <= TOL


@pytest.mark.parametrize("mode,tol", [("mfma", 1e-5), ("fp64", 1e-6)])
@pytest.mark.parametrize("which", ["small", "clinical"])
def test_ramp_filter_matches_oracle(which, mode, tol, request):
    # mfma: BF16 matrix cores with three-way split operands and FP32 accumulation
    # (FP32-GEMM-level error); fp64: FP64 products and sums (only the output rounding)
    plan, o, img = request.getfixturevalue(which)
    for kind in ("sinogram", "noise"):
        if kind == "sinogram":
            sino = o.fp(img).astype(np.float32)
        else:
            sino = np.random.default_rng(11).normal(size=(2,) + plan.sino_shape).astype(np.float32)
        got = host(plan.ramp_filter(dev(sino), mode=mode))
        err = rel_l2(got, o.ramp_filter(sino))
        _say("ramp %s %s %s rel_l2=%.3e" % (mode, which, kind, err))
        assert err <= tol


def test_fdk_equals_fbp(small):
    plan, o, img = small
    sino = dev(o.fp(img))
    a = plan.fdk(sino)
    b = plan.fbp(sino)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


# ------------------------------------------------------------------- adjoints

def test_fp_adjoint_matches_oracle(small):
    plan, o, img = small
    y = np.random.default_rng(12).normal(size=(1, 360, 512)).astype(np.float32)
    got = host(plan.fp_adjoint(dev(y)))
    err = rel_l2(got, o.fp_adjoint(y))
    _say("fp_adjoint rel_l2=%.3e" % err)
    assert err <= TOL


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("image", ["phantom", "white_noise"])
@pytest.mark.parametrize("which", ["small", "clinical"])
def test_bp_adjoint_matches_oracle(which, weighted, image, request):
    # BP^T and FBP^T gather per (view, column) with the column offset taken relative to the
    # column's node ray (a few 1e-7 column), so white noise, where every interpolation
    # weight counts independently, and FBP^T, whose ramp filter amplifies the rounding of
    # those weights, stay at the 1e-5 bar too
    plan, o, img = request.getfixturevalue(which)
    if image == "phantom":
        x = img.astype(np.float32)
    else:
        x = np.random.default_rng(13).normal(size=(1,) + plan.image_shape).astype(np.float32)
    if weighted:
        got = host(plan.fbp_adjoint(dev(x)))
        ref = o.fbp_adjoint(x)
    else:
        got = host(plan.bp_adjoint(dev(x)))
        ref = o.bp_adjoint(x)
    err = rel_l2(got, ref)
    _say("bp_adjoint %s weighted=%s %s rel_l2=%.3e" % (which, weighted, image, err))
    assert err <= TOL


@pytest.mark.parametrize("n", [1, 5, 16, 19])
def test_bp_adjoint_deterministic_and_slice_independent(clinical, n):
    # a fixed summation order per (view, column): repeated calls are bitwise equal, and a
    # slice's result does not depend on the slice group it runs in (16-, 8-, 4-, 2-slice
    # groups and single slices share the arithmetic)
    plan, o, img = clinical
    x = np.random.default_rng(n).normal(size=(n,) + plan.image_shape).astype(np.float32)
    a1, a2 = plan.bp_adjoint(dev(x)), plan.bp_adjoint(dev(x))
    f1, f2 = plan.fbp_adjoint(dev(x)), plan.fbp_adjoint(dev(x))
    torch.cuda.synchronize()
    assert torch.equal(a1, a2) and torch.equal(f1, f2)
    for s in (0, n // 2, n - 1):
        assert torch.equal(a1[s], plan.bp_adjoint(dev(x[s:s + 1]))[0])
        assert torch.equal(f1[s], plan.fbp_adjoint(dev(x[s:s + 1]))[0])


def _dot(a, b):
    return float(np.dot(np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()))


@pytest.mark.parametrize("op", ["fp", "bp", "fbp"])
def test_adjoint_dot_product_clinical(clinical, op):
    plan, o, img = clinical
    rng = np.random.default_rng(14)
    x = dev(rng.normal(size=(1, 512, 512)))
    y = dev(rng.normal(size=(1, 984, 888)))
    # <A u, w> vs <u, A^T w>, normalised by the Cauchy-Schwarz bound |A u| |w| (random
    # u, w make the raw dot product a sum with heavy cancellation)
    if op == "fp":
        au, atw = host(plan.fp(x)), host(plan.fp_adjoint(y))
        lhs, rhs, scale = _dot(au, host(y)), _dot(host(x), atw), np.linalg.norm(au) * np.linalg.norm(host(y))
    else:
        fwd = plan.bp if op == "bp" else plan.fbp
        adj = plan.bp_adjoint if op == "bp" else plan.fbp_adjoint
        au, atw = host(fwd(y)), host(adj(x))
        lhs, rhs, scale = _dot(au, host(x)), _dot(host(y), atw), np.linalg.norm(au) * np.linalg.norm(host(x))
    rel = abs(lhs - rhs) / scale
    _say("dot %s: %.6e vs %.6e rel=%.2e (of |Au||w|)" % (op, lhs, rhs, rel))
    assert rel <= TOL


# ---------------------------------------------------------------- edge cases

@pytest.mark.parametrize("n", [2, 3, 5, 8])
def test_batch_equals_per_slice(small, n):
    # batches share each sample's position across slices (a different but equivalent
    # bilinear arithmetic), so slices agree with single-slice calls to FP32 rounding
    plan, o, img = small
    x = np.stack([np.roll(img[0], 17 * s, axis=1) for s in range(n)])
    y = np.random.default_rng(n).normal(size=(n, 360, 512)).astype(np.float32)
    bf, ba = host(plan.fp(dev(x))), host(plan.fp_adjoint(dev(y)))
    for s in range(n):
        assert rel_l2(bf[s], host(plan.fp(dev(x[s:s + 1])))[0]) <= 1e-6
        assert rel_l2(ba[s], host(plan.fp_adjoint(dev(y[s:s + 1])))[0]) <= 1e-6


@pytest.mark.parametrize("n", [2, 3, 4, 8, 12, 16, 19])
def test_clinical_batch_equals_per_slice(clinical, n):
    # the clinical grid selects the shared-slice kernels: FP with 2-slice tiles (n = 2, 3:
    # two tile classes, 88 of the 128 rows) or 4-slice tiles, the
    # adjoint with 16-, 8- or 4-slice groups staged through LDS (one slice: the
    # cell-driven kernel); a ragged last group is padded.  The adjoint runs on smooth
    # sinograms (the slices' projections): the gather and cell-driven kernels round the
    # FP32 sample positions differently, which on white noise alone differs by about
    # 1e-5 relative (each is within 2e-5 of the FP64 oracle there)
    plan, o, img = clinical
    x = np.stack([np.roll(img[0], 23 * s, axis=1) for s in range(n)])
    bf = host(plan.fp(dev(x)))
    y = bf.astype(np.float32)
    ba = host(plan.fp_adjoint(dev(y)))
    for s in (0, n // 2, n - 1):
        assert rel_l2(bf[s], host(plan.fp(dev(x[s:s + 1])))[0]) <= 1e-6
        assert rel_l2(ba[s], host(plan.fp_adjoint(dev(y[s:s + 1])))[0]) <= 1e-6


@pytest.mark.parametrize("n", [8, 16, 32, 37])
def test_clinical_bp_groups_equal_per_slice(clinical, n):
    # on the clinical grid BP and FBP run slice groups of 16, 8 or 4 that stage their
    # detector columns in LDS (37 = 2 x 16 + 4 + 1: a staged group and an unstaged slice)
    plan, o, img = clinical
    y = np.random.default_rng(n).normal(size=(n, 984, 888)).astype(np.float32)
    bb, bf = host(plan.bp(dev(y))), host(plan.fbp(dev(y)))
    for s in (0, n // 2, n - 1):
        assert rel_l2(bb[s], host(plan.bp(dev(y[s:s + 1])))[0]) <= 1e-6
        assert rel_l2(bf[s], host(plan.fbp(dev(y[s:s + 1])))[0]) <= 1e-6


def test_staged_groups_on_ragged_tiles(oracle_mod):
    # 500 x 468 pixels (neither side a multiple of the 16-pixel tile) and 48 slices: BP and
    # FBP run three 16-slice groups that stage their detector columns in LDS, with edge
    # tiles partly outside the image; each slice equals its single-slice call (the
    # unstaged kernel) and the first, a phantom's projections (the oracle tolerance is
    # stated for phantoms), matches the oracle
    p, a = workloads.scan(nx=500, ny=468, rfov=250.0, sid=1000.0, sdd=2000.0, n_cols=256, dcol_rad=0.0022,
                          n_views=180)
    plan = FanBeamPlan(p, a)
    n = 48
    y = np.random.default_rng(48).uniform(0, 1, size=(n,) + plan.sino_shape).astype(np.float32)
    y[0] = host(plan.fp(dev(workloads.ellipse_phantom(500, 468, 250.0, 8, seed=4)[None])))[0]
    bb, bf = host(plan.bp(dev(y))), host(plan.fbp(dev(y)))
    for s in (0, n // 2, n - 1):
        assert rel_l2(bb[s], host(plan.bp(dev(y[s:s + 1])))[0]) <= 1e-6
        assert rel_l2(bf[s], host(plan.fbp(dev(y[s:s + 1])))[0]) <= 1e-6
    o = oracle_mod.Oracle(p, a)
    for name, got, ref in (("bp", bb[:1], o.bp(y[:1])), ("fbp", bf[:1], o.fbp(y[:1]))):
        err = rel_l2(got, ref)
        _say("ragged 16-slice groups %s rel_l2=%.3e" % (name, err))
        assert err <= TOL, (name, err)


def test_zero_slices():
    plan = FanBeamPlan(*workloads.tiny_scan())
    out = plan.fp(torch.zeros(0, 16, 16, device="cuda"))
    assert out.shape == (0, 24, 20)


@pytest.mark.parametrize("case", ["wide_fan", "shift", "nonsquare", "odd_sizes", "few_views"])
def test_edge_geometries(oracle_mod, case):
    if case == "wide_fan":      # outer rays miss the FOV; FOV edge pixels leave the fan
        p, a = workloads.tiny_scan(nx=48, n_views=40, n_cols=70, rfov=24.0, sid=100.0, fan_margin=1.6)
    elif case == "shift":       # detector shift: FP ignores it, BP/FBP apply it
        p, a = workloads.tiny_scan(nx=40, n_views=36, n_cols=64, rfov=20.0, sid=90.0, dshift=1.125)
    elif case == "nonsquare":   # nX != nY (the reference scales both axes by dx in FP)
        p, a = workloads.scan(nx=48, ny=40, rfov=24.0, sid=100.0, sdd=200.0, n_cols=80, dcol_rad=0.006,
                              n_views=30)
    elif case == "odd_sizes":
        p, a = workloads.tiny_scan(nx=37, n_views=29, n_cols=53, rfov=18.5, sid=70.0)
    else:
        p, a = workloads.tiny_scan(nx=64, n_views=3, n_cols=97, rfov=32.0, sid=150.0)
    plan = FanBeamPlan(p, a)
    o = oracle_mod.Oracle(p, a)
    rng = np.random.default_rng(15)
    x = rng.uniform(0, 1, size=(2,) + plan.image_shape).astype(np.float32)
    y = rng.uniform(0, 1, size=(2,) + plan.sino_shape).astype(np.float32)
    for name, got, ref in (
        ("fp", plan.fp(dev(x)), o.fp(x)),
        ("bp", plan.bp(dev(y)), o.bp(y)),
        ("fbp", plan.fbp(dev(y)), o.fbp(y)),
        ("fp_adjoint", plan.fp_adjoint(dev(y)), o.fp_adjoint(y)),
        ("bp_adjoint", plan.bp_adjoint(dev(x)), o.bp_adjoint(x)),
        ("fbp_adjoint", plan.fbp_adjoint(dev(x)), o.fbp_adjoint(x)),
    ):
        err = rel_l2(host(got), ref)
        _say("%s %s rel_l2=%.3e" % (case, name, err))
        assert err <= TOL, (case, name, err)


# ------------------------------------------------------------ drop-in surface

def test_dropin_compat_layer_roundtrip(oracle_mod):
    # the reference's own call pattern (utils.py) with its config.py geometry, 2 images
    img = np.stack([workloads.ellipse_phantom(512, 512, 200.0, 8, seed=s) for s in (21, 22)])
    sino = compat.ForwardProjection(dev(img), nimg=2)
    assert tuple(sino.shape) == (2, 984, 888)
    o = oracle_mod.Oracle(compat.param_set(2), compat.angle(984, 0))
    assert rel_l2(host(sino), o.fp(img)) <= TOL
    rec = compat.FilteredBackProjection(sino, nimg=2)
    assert tuple(rec.shape) == (2, 512, 512)
    bp = compat.BackProjection(sino, nimg=2)
    assert rel_l2(host(bp), o.bp(host(sino).astype(np.float32))) <= TOL
    with pytest.raises(ValueError, match="parameter\\[11\\]"):
        compat.ForwardProjection(dev(img), nimg=3)


def test_dropin_fdk_op_all_slices(small):
    plan, o, img = small
    sino = o.fp(img).astype(np.float32)
    two = np.concatenate([sino, 0.5 * sino])
    out = ops.BackProjection(dev(two), plan.angle, plan.parameter, [2, 256, 256], [2, 360, 512])
    ref = o.fbp(two)
    assert rel_l2(host(out), ref) <= TOL


# ------------------------------------------------------------------- autograd

def test_fp_gradient_matches_oracle(small):
    plan, o, img = small
    rng = np.random.default_rng(16)
    target = (o.fp(img) + rng.normal(0, 0.01, size=(1, 360, 512))).astype(np.float32)
    x0 = (img * 0.8).astype(np.float32)
    x = dev(x0).requires_grad_(True)
    loss = ((ops.forward_project(x, plan) - dev(target)) ** 2).sum()
    loss.backward()
    g = host(x.grad)
    g_ref = 2.0 * o.fp_adjoint(o.fp(x0) - target)
    err = rel_l2(g, g_ref)
    _say("fp grad rel_l2=%.3e" % err)
    assert err <= 1e-4  # residual y - Ax is formed in FP32 before the adjoint


# -------------------------------------------------------------- BF16 storage

def _bf16_round(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16).float().numpy()


def test_fp_bf16_storage(clinical):
    # BF16 image storage, FP32 arithmetic: exact against the oracle on the BF16-rounded
    # image, and within 1e-3 of the FP32 image's projection
    plan, o, img = clinical
    xb = torch.from_numpy(img).cuda().to(torch.bfloat16)
    got = host(plan.fp(xb))
    e_storage = rel_l2(got, o.fp(_bf16_round(img)))
    e_total = rel_l2(got, o.fp(img))
    _say("fp bf16-in: vs rounded input %.3e, vs fp32 input %.3e" % (e_storage, e_total))
    assert e_storage <= TOL and e_total <= 1e-3
    out_b = plan.fp(torch.from_numpy(img).cuda(), out_dtype=torch.bfloat16)
    assert out_b.dtype == torch.bfloat16
    e_out = rel_l2(host(out_b.float()), o.fp(img))
    _say("fp bf16-out: %.3e" % e_out)
    assert e_out <= 4e-3  # one BF16 rounding of the output (2^-9 relative)


def test_fp_adjoint_bf16_storage(clinical):
    plan, o, img = clinical
    y = (o.fp(img) * (1 + np.random.default_rng(17).normal(0, 0.01, size=(1, 984, 888)))).astype(np.float32)
    got = host(plan.fp_adjoint(torch.from_numpy(y).cuda().to(torch.bfloat16)))
    e_storage = rel_l2(got, o.fp_adjoint(_bf16_round(y)))
    e_total = rel_l2(got, o.fp_adjoint(y))
    _say("fp_adjoint bf16-in: vs rounded input %.3e, vs fp32 input %.3e" % (e_storage, e_total))
    assert e_storage <= TOL and e_total <= 1e-3


def test_bf16_image_gradient_dtype(small):
    plan, o, img = small
    x = torch.from_numpy(img).cuda().to(torch.bfloat16).requires_grad_(True)
    y = ops.forward_project(x, plan)
    assert y.dtype == torch.float32
    (y ** 2).sum().backward()
    assert x.grad.dtype == torch.bfloat16
    g_ref = 2.0 * o.fp_adjoint(o.fp(_bf16_round(img)))
    err = rel_l2(host(x.grad.float()), g_ref)
    _say("bf16 grad rel_l2=%.3e" % err)
    assert err <= 4e-3
