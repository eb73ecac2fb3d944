"""GPU, full workload sizes, through size-independent properties and sampled oracle
checks (the FP64 oracle cannot process whole volumes in seconds).

* Cone FDK at 512x512x256 from 720 views of a 1024x256 panel: 48 filtered lines against
  the oracle's taps and pre-weights (1e-5), the backprojection of the GPU-filtered
  projections on two slices against the oracle (1e-5), and linearity of the whole FDK
  (1e-5 relative to the result: each term carries FP32 rounding of ~1e-6).
* Uneven batch of 24 slices drawn like the 256-slice workload (sides up to 768, up to
  1200 views, up to 1024 bins): every slice equals its single-plan call bit for bit, the
  batch FP is linear, and the batch pair passes the adjoint dot test (1e-5).
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from fanbeamctrecon_tensorflowop_amd import workloads
from fanbeamctrecon_tensorflowop_amd.batch import UnevenBatch
from fanbeamctrecon_tensorflowop_amd.cone import ConePlan
from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_fdk_full_size_sampled_oracle_and_linearity(oracle_mod):
    g, a = workloads.cone_scan()
    plan = ConePlan(g, a)
    ells = workloads.random_ellipsoids(10, seed=5)
    proj = workloads.ellipsoid_projections(ells, g, a, device="cuda").float()
    q = plan.filter(proj[None])[0]
    o = oracle_mod.ConeOracle(g, a)
    h, w = o.taps(), o.preweight()
    rng = np.random.default_rng(1)
    nC = g["nCols"]
    num = den = 0.0
    for _ in range(48):
        v, r = int(rng.integers(0, g["nViews"])), int(rng.integers(0, g["nRows"]))
        line = proj[v, r].double().cpu().numpy()
        ref = np.convolve(line * w[r], h)[nC - 1:2 * nC - 1]
        got = q[v, r].double().cpu().numpy()
        num += float(((got - ref) ** 2).sum())
        den += float((ref ** 2).sum())
    e_f = (num / den) ** 0.5
    vol = plan.backproject(q[None])[0]
    qh = q.double().cpu().numpy()
    e_b = []
    for k in (60, 190):
        gk = dict(g, nZ=1, zmin=float(np.float32(g["zmin"]) + np.float32(k) * np.float32(g["dz"])))
        ref = oracle_mod.ConeOracle(gk, a).backproject(qh[None])[0, 0]
        e_b.append(rel_l2(vol[k].double().cpu().numpy(), ref))
    _say("fdk full size: filter lines %.3e, backprojection slices %s" % (e_f, ["%.3e" % e for e in e_b]))
    assert e_f <= 1e-5 and max(e_b) <= 1e-5
    # linearity of the whole FDK
    p2 = torch.roll(proj, 37, dims=2)
    lhs = plan.fdk((0.75 * proj - 0.5 * p2)[None])[0]
    rhs = 0.75 * plan.fdk(proj[None])[0] - 0.5 * plan.fdk(p2[None])[0]
    e_l = float(torch.linalg.vector_norm((lhs - rhs).double()) / torch.linalg.vector_norm(rhs.double()))
    _say("fdk full size linearity %.3e" % e_l)
    assert e_l <= 1e-5


def test_uneven_batch_full_sizes_properties():
    scans = workloads.uneven_scans(24, seed=4)
    plans = [FanBeamPlan(p, a) for p, a in scans]
    batch = UnevenBatch(plans)
    batch.set_shape(fp_rays=64, fp_wgs_per_cu=8, adj_tile_x=16, adj_wgs_per_cu=8)
    x = batch.pack_images(workloads.uneven_phantoms(scans, seed=4))
    y = batch.fp(x)
    r = torch.from_numpy(np.random.default_rng(2).normal(size=batch.sino_numel).astype(np.float32)).cuda()
    g = batch.fp_adjoint(r)
    for s in range(len(plans)):
        assert torch.equal(batch.sino(y, s), plans[s].fp(batch.image(x, s)[None])[0]), "fp slice %d" % s
        assert torch.equal(batch.image(g, s), plans[s].fp_adjoint(batch.sino(r, s)[None])[0]), "adjoint slice %d" % s
    x2 = torch.roll(x, 11)
    lin = batch.fp(2.0 * x - 0.25 * x2)
    e_l = rel_l2(lin.double().cpu().numpy(), (2.0 * y - 0.25 * batch.fp(x2)).double().cpu().numpy())
    lhs = float(torch.dot(y.double(), r.double()))
    rhs = float(torch.dot(x.double(), g.double()))
    e_d = abs(lhs - rhs) / (float(torch.linalg.vector_norm(y.double())) * float(torch.linalg.vector_norm(r.double())))
    _say("uneven full sizes: linearity %.3e, dot %.3e" % (e_l, e_d))
    assert e_l <= 1e-6 and e_d <= 1e-5
