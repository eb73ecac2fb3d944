"""GPU: the full 256-slice uneven batch (seed 4: image side in {256, 384, 512, 768}, 360-1200
views, 512-1024 bins, one FOV and fan), FP32 and BF16 storage, with the autotuned launch
shape, through size-independent properties:

  * every slice of the batch FP and FP^T equals its single-plan call (bit for bit when the
    tuned shape runs the single-plan kernel body, 64-ray bundles and 16x16 tiles; else
    within 1e-6 relative L2, the same per-sample arithmetic on another tiling);
  * the batch FP is linear (1e-6) and the batch pair passes the adjoint dot test
    <FP x, r> = <x, FP^T r> at 1e-5 of |FP x||r|;
  * BF16 storage: each slice equals the single-plan BF16 path bit for bit, and its FP
    stays within one BF16 rounding (2e-3) of the FP32-storage FP.
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from fanbeamctrecon_tensorflowop_amd import workloads
from fanbeamctrecon_tensorflowop_amd.batch import UnevenBatch, autotune
from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan

pytestmark = pytest.mark.gpu

N_SLICES = 256


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c4(tmp_path_factory):
    scans = workloads.uneven_scans(N_SLICES, seed=4)
    plans = [FanBeamPlan(p, a) for p, a in scans]
    batch = UnevenBatch(plans)
    phantoms = workloads.uneven_phantoms(scans, seed=4)
    _say("c4: %d slices, %.3g updates, %.3g FP samples" % (len(plans), batch.updates, batch.fp_samples))
    return plans, batch, phantoms, str(tmp_path_factory.mktemp("tune") / "shapes.json")


def _same_kernel_body(shape):
    return shape["fp_rays"] == 64 and shape["adj_tile_x"] == 16


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_c4_full_batch_autotuned(c4, storage, monkeypatch):
    plans, batch, phantoms, cache = c4
    monkeypatch.setenv("FBCT_TUNE_CACHE", cache)
    dt = torch.float32 if storage == "f32" else torch.bfloat16
    x = batch.pack_images(phantoms, dtype=dt)
    y = batch.fp(x)
    tuned = autotune(batch, x, y.to(dt), reps=1)
    shape = batch.shape
    _say("c4 %s autotuned %s (fp %.2f ms, adj %.2f ms)" % (storage, shape, tuned["fp_ms"], tuned["adj_ms"]))
    y = batch.fp(x)
    r = torch.from_numpy(np.random.default_rng(3).normal(size=batch.sino_numel).astype(np.float32)).cuda()
    rs = r.to(dt)
    g = batch.fp_adjoint(rs)
    torch.cuda.synchronize()
    exact = _same_kernel_body(shape)
    worst_f = worst_a = 0.0
    for s, p in enumerate(plans):
        ys = p.fp(batch.image(x, s)[None])[0]
        gs = p.fp_adjoint(batch.sino(rs, s)[None])[0]
        if exact:
            assert torch.equal(batch.sino(y, s), ys), "fp slice %d" % s
            assert torch.equal(batch.image(g, s), gs), "adjoint slice %d" % s
        else:
            worst_f = max(worst_f, rel_l2(batch.sino(y, s).double().cpu().numpy(), ys.double().cpu().numpy()))
            worst_a = max(worst_a, rel_l2(batch.image(g, s).double().cpu().numpy(), gs.double().cpu().numpy()))
    _say("c4 %s per-slice: %s (worst fp %.2e, adj %.2e)" % (storage, "bitwise" if exact else "rel", worst_f, worst_a))
    assert worst_f <= 1e-6 and worst_a <= 1e-6
    # linearity and the adjoint dot test on the whole batch (FP32 arithmetic on the
    # storage-rounded inputs)
    xf = x.float()
    x2 = torch.roll(xf, 11)
    lin = batch.fp(2.0 * xf - 0.25 * x2)
    e_l = rel_l2(lin.double().cpu().numpy(), (2.0 * batch.fp(xf) - 0.25 * batch.fp(x2)).double().cpu().numpy())
    lhs = float(torch.dot(y.double(), rs.double()))
    rhs = float(torch.dot(xf.double(), g.double()))
    e_d = abs(lhs - rhs) / (float(torch.linalg.vector_norm(y.double())) * float(torch.linalg.vector_norm(rs.double())))
    _say("c4 %s: linearity %.3e, dot %.3e" % (storage, e_l, e_d))
    assert e_l <= 1e-6 and e_d <= 1e-5
    if storage == "bf16":
        # BF16 storage against FP32 storage of the same phantoms: one BF16 rounding of the
        # image (a relative ulp of 2^-8 to 2^-7); the ellipse phantoms are piecewise
        # constant, so the rounding is the same over each ellipse and does not average out
        # along a ray (1.04e-3 measured)
        xf32 = batch.pack_images(phantoms)
        yf32 = batch.fp(xf32)
        e_y = rel_l2(y.double().cpu().numpy(), yf32.double().cpu().numpy())
        _say("c4 bf16 fp vs fp32 storage %.3e" % e_y)
        assert e_y <= 2e-3
