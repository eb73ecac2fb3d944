"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
FP64 oracle): the oracle must reproduce them (CPU), and the GPU operators must match
them (GPU) at the parity tolerances:
  fan: FP, FP^T, BP, ramp filter, FBP, BP^T, FBP^T 1e-5; cone flat FDK 1e-5."""
import json
import os

import numpy as np
import pytest

from conftest import rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(GOLD, name)) as z:  # allow_pickle=False (default)
        return {k: z[k] for k in z.files}


def test_oracle_reproduces_fan_fixture(oracle_mod):
    d = _load("fan_48px_60v.npz")
    o = oracle_mod.Oracle(d["parameter"], d["angles"])
    img, sino = d["image"].astype(np.float64), d["sino"].astype(np.float64)
    for name, got in (("fp", o.fp(img)), ("fp_adjoint", o.fp_adjoint(sino)), ("bp", o.bp(sino)),
                      ("bp_adjoint", o.bp_adjoint(img)), ("fbp", o.fbp(sino)), ("fbp_adjoint", o.fbp_adjoint(img)),
                      ("ramp", o.ramp_filter(sino))):
        np.testing.assert_allclose(got, d[name], rtol=1e-12, atol=1e-12 * np.abs(d[name]).max(), err_msg=name)


def test_oracle_reproduces_cone_fixture(oracle_mod):
    d = _load("cone_flat_32px_48v.npz")
    g = json.loads(str(d["geom"]))
    got = oracle_mod.ConeOracle(g, d["angles"]).fdk(d["proj"][None].astype(np.float64))[0]
    np.testing.assert_allclose(got, d["fdk"], rtol=1e-12, atol=1e-12 * np.abs(d["fdk"]).max())


@pytest.mark.gpu
def test_gpu_matches_fan_fixture():
    import torch
    from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan
    d = _load("fan_48px_60v.npz")
    plan = FanBeamPlan(d["parameter"], d["angles"])
    img = torch.from_numpy(d["image"]).cuda()
    sino = torch.from_numpy(d["sino"]).cuda()
    cases = (("fp", plan.fp(img), 1e-5), ("fp_adjoint", plan.fp_adjoint(sino), 1e-5), ("bp", plan.bp(sino), 1e-5),
             ("ramp", plan.ramp_filter(sino), 1e-5), ("fbp", plan.fbp(sino), 1e-5),
             ("bp_adjoint", plan.bp_adjoint(img), 1e-5), ("fbp_adjoint", plan.fbp_adjoint(img), 1e-5))
    for name, got, tol in cases:
        e = rel_l2(got.double().cpu().numpy(), d[name])
        print("golden fan %s rel_l2=%.3e" % (name, e), flush=True)
        assert e <= tol, name


@pytest.mark.gpu
def test_gpu_matches_cone_fixture():
    import torch
    from fanbeamctrecon_tensorflowop_amd.cone import ConePlan
    d = _load("cone_flat_32px_48v.npz")
    plan = ConePlan(json.loads(str(d["geom"])), d["angles"])
    got = plan.fdk(torch.from_numpy(d["proj"]).cuda())[0].double().cpu().numpy()
    e = rel_l2(got, d["fdk"])
    print("golden cone fdk rel_l2=%.3e" % e, flush=True)
    assert e <= 1e-5
