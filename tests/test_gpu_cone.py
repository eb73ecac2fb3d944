"""GPU: cone-beam FDK (flat and curved detectors, multi-row, multi-slice) against the
FP64 oracle.

Tolerances (relative L2):
  FP32 projections, FP32 volume .................... 1e-5 against the oracle
  BF16 projections (FP32 arithmetic) ............... 1e-5 against the oracle fed the same
                                                     BF16-rounded projections; the error
                                                     against the FP32 projections is printed
                                                     and bounded by 1e-2: the ramp filter
                                                     amplifies the white BF16 rounding noise
                                                     (2^-9 relative per sample; 5.1e-3 measured
                                                     on this case)
  BF16 volume output ............................... 4e-3 against the FP32 output (2^-9 rounding)
  curved, nZ = nRows = 1 ........................... 1e-6 against the fan-beam FBP kernels
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from fanbeamctrecon_tensorflowop_amd import workloads
from fanbeamctrecon_tensorflowop_amd.cone import ConePlan
from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def flat_case(oracle_mod):
    g, a = workloads.cone_scan(nx=128, nz=64, dx=0.6, n_views=180, n_cols=256, n_rows=96, pitch=0.8,
                               R=300.0, D=560.0)
    ells = workloads.random_ellipsoids(8, seed=21, rfov=38.4, zhalf=19.2)
    P = workloads.ellipsoid_projections(ells, g, a).numpy()
    P += np.random.default_rng(3).normal(0, 1e-3, size=P.shape)
    P = P.astype(np.float32)
    o = oracle_mod.ConeOracle(g, a)
    ref = o.fdk(P[None].astype(np.float64))[0]
    return g, a, P, o, ref


def test_flat_fdk_matches_oracle(flat_case):
    g, a, P, o, ref = flat_case
    plan = ConePlan(g, a)
    got = plan.fdk(torch.from_numpy(P).cuda())[0].double().cpu().numpy()
    e = rel_l2(got, ref)
    _say("flat fdk rel_l2=%.3e" % e)
    assert e <= 1e-5


def test_flat_filter_matches_oracle(flat_case):
    g, a, P, o, ref = flat_case
    plan = ConePlan(g, a)
    q = plan.filter(torch.from_numpy(P).cuda())[0].double().cpu().numpy()
    e = rel_l2(q, o.filter(P[None].astype(np.float64))[0])
    _say("flat filter rel_l2=%.3e" % e)
    assert e <= 1e-5


def test_flat_bf16_projections(flat_case):
    g, a, P, o, ref = flat_case
    plan = ConePlan(g, a)
    pb = torch.from_numpy(P).cuda().to(torch.bfloat16)
    got = plan.fdk(pb)[0].double().cpu().numpy()
    ref_b = o.fdk(pb.float().cpu().numpy()[None].astype(np.float64))[0]
    e_same = rel_l2(got, ref_b)
    e_fp32 = rel_l2(got, ref)
    _say("flat fdk bf16 projections: vs oracle(bf16 inputs) %.3e, vs oracle(fp32 inputs) %.3e" % (e_same, e_fp32))
    assert e_same <= 1e-5
    assert e_fp32 <= 1e-2


def test_flat_bf16_volume_output(flat_case):
    g, a, P, o, ref = flat_case
    plan = ConePlan(g, a)
    x = torch.from_numpy(P).cuda()
    v32 = plan.fdk(x)
    v16 = plan.fdk(x, out_dtype=torch.bfloat16)
    e = rel_l2(v16.double().cpu().numpy(), v32.double().cpu().numpy())
    _say("bf16 volume output rel_l2=%.3e" % e)
    assert e <= 4e-3


@pytest.mark.parametrize("n_rows,n_z", [(7, 5), (24, 16)])
def test_curved_multirow_matches_oracle(oracle_mod, n_rows, n_z):
    p, a = workloads.scan(nx=96, rfov=48.0, sid=250.0, sdd=480.0, n_cols=160, dcol_rad=0.44 / 160, n_views=150)
    p = p.copy()
    p[13], p[14], p[15] = -0.5 * n_rows * 1.1, 1.1, 0.9  # detBottom, dRow, dz
    plan = ConePlan.from_params(p, a, n_rows, n_z)
    ells = workloads.random_ellipsoids(6, seed=n_rows, rfov=48.0, zhalf=0.5 * n_z * 0.9)
    proj = workloads.ellipsoid_projections(ells, plan.geom, a).numpy()
    proj = np.stack([proj, proj[:, ::-1, :]]) + np.random.default_rng(n_rows).normal(0, 1e-3, (2,) + proj.shape)
    proj = proj.astype(np.float32)
    got = plan.fdk(torch.from_numpy(proj).cuda()).double().cpu().numpy()
    ref = oracle_mod.Oracle(p, a).fdk(proj.astype(np.float64), n_rows=n_rows, n_z=n_z)
    e = rel_l2(got, ref)
    _say("curved rows=%d z=%d fdk rel_l2=%.3e" % (n_rows, n_z, e))
    assert e <= 1e-5


@pytest.mark.parametrize("detector", ["flat", "curved"])
def test_fine_pitch_fdk_matches_oracle(oracle_mod, detector):
    """Detector cells 4x finer than the magnified voxels: a 16 x 16 x 16 voxel tile's
    footprint (about 120 x 120 samples per view) exceeds the 64-column, 44-row LDS staging
    window, so most columns take their taps from global memory and the rest from LDS."""
    g, a = workloads.cone_scan(nx=64, nz=32, dx=1.0, n_views=90, n_cols=640, n_rows=280, pitch=0.25,
                               R=150.0, D=300.0, detector=detector)
    ells = workloads.random_ellipsoids(6, seed=33, rfov=32.0, zhalf=16.0)
    P = workloads.ellipsoid_projections(ells, g, a).numpy()
    P = (P + np.random.default_rng(9).normal(0, 1e-3, size=P.shape)).astype(np.float32)
    ref = oracle_mod.ConeOracle(g, a).fdk(P[None].astype(np.float64))[0]
    got = ConePlan(g, a).fdk(torch.from_numpy(P).cuda())[0].double().cpu().numpy()
    e = rel_l2(got, ref)
    _say("%s fine-pitch fdk rel_l2=%.3e" % (detector, e))
    assert e <= 1e-5


def test_curved_single_row_equals_fan_fbp():
    p, a = workloads.clinical_scan(3)
    plan = ConePlan.from_params(p, a, 1, 1)
    fan = FanBeamPlan(p, a)
    sino = fan.fp(torch.from_numpy(workloads.clinical_phantom(3)).cuda())
    vol = plan.fdk(sino[:, :, None, :])[:, 0]
    e = rel_l2(vol.double().cpu().numpy(), fan.fbp(sino).double().cpu().numpy())
    _say("curved nZ=nRows=1 vs fan fbp rel_l2=%.3e" % e)
    assert e <= 1e-6


def test_volumes_batch_equals_single(flat_case):
    g, a, P, o, ref = flat_case
    plan = ConePlan(g, a)
    x = torch.from_numpy(np.stack([P, 0.5 * P[::-1].copy()])).cuda()
    both = plan.fdk(x)
    for k in range(2):
        one = plan.fdk(x[k:k + 1])
        assert rel_l2(both[k].double().cpu().numpy(), one[0].double().cpu().numpy()) <= 1e-6


def test_ragged_tiles_match_oracle(oracle_mod):
    # 88 x 88 x 16 voxels: the LDS-staged backprojection (16-voxel z-chunks) with edge tiles
    # half outside the volume
    g, a = workloads.cone_scan(nx=88, nz=16, dx=0.6, n_views=90, n_cols=192, n_rows=48, pitch=0.8, R=300.0,
                               D=560.0)
    ells = workloads.random_ellipsoids(6, seed=5, rfov=26.4, zhalf=4.0)
    P = workloads.ellipsoid_projections(ells, g, a).numpy().astype(np.float32)
    ref = oracle_mod.ConeOracle(g, a).fdk(P[None].astype(np.float64))[0]
    got = ConePlan(g, a).fdk(torch.from_numpy(P).cuda())[0].double().cpu().numpy()
    e = rel_l2(got, ref)
    _say("ragged cone fdk rel_l2=%.3e" % e)
    assert e <= 1e-5


def test_filter_then_backproject_equals_fdk(flat_case):
    # fbct_cone_fdk reads column pairs written by the filter; the separate entry points
    # use the plain layout: both give the same volume
    g, a, P, o, ref = flat_case
    plan = ConePlan(g, a)
    x = torch.from_numpy(P).cuda()
    two_step = plan.backproject(plan.filter(x))
    one_step = plan.fdk(x)
    assert rel_l2(two_step.double().cpu().numpy(), one_step.double().cpu().numpy()) <= 1e-6
