"""GPU: the reference ops as torch custom ops (torch.ops.fbct.*).

  * each op equals the plan-level call on the same data (bit for bit: same kernels, each
    sum in a fixed order);
  * gradients through the ops equal the matched adjoints (bit for bit);
  * torch.library.opcheck: schema, autograd registration and fake-tensor propagation.
"""
import numpy as np
import pytest
import torch


from fanbeamctrecon_tensorflowop_amd import ops, workloads
from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    p, a = workloads.scan(nx=64, rfov=32.0, sid=100.0, sdd=200.0, n_cols=96, dcol_rad=0.7 / 96, n_views=72,
                          n_slices=3)
    plan = FanBeamPlan(p, a)
    img = torch.from_numpy(np.stack([workloads.ellipse_phantom(64, 64, 32.0, 8, s) for s in range(3)])).cuda()
    vol, proj = [3, 64, 64], [3, 72, 96]
    return p, a, plan, img, vol, proj


def test_ops_equal_plan_calls(case):
    p, a, plan, img, vol, proj = case
    al, pl = a.tolist(), p.tolist()
    sino = plan.fp(img)
    assert torch.equal(torch.ops.fbct.fp_fan_curve(img, al, pl, vol, proj), sino)
    assert torch.equal(torch.ops.fbct.fp_fan_curve_adjoint(sino, al, pl, vol, proj), plan.fp_adjoint(sino))
    assert torch.equal(torch.ops.fbct.bp_fan_curve(sino, al, pl, vol, proj), plan.bp(sino))
    assert torch.equal(torch.ops.fbct.bp_fan_curve_adjoint(img, al, pl, vol, proj), plan.bp_adjoint(img))
    assert torch.equal(torch.ops.fbct.fbp_fan_curve(sino, al, pl, vol, proj), plan.fbp(sino))
    assert torch.equal(torch.ops.fbct.fbp_fan_curve_adjoint(img, al, pl, vol, proj), plan.fbp_adjoint(img))
    assert torch.equal(torch.ops.fbct.back_projection(sino, al, pl, vol, proj), plan.fdk(sino))
    # the reference-keyword Python shims dispatch to the same ops
    assert torch.equal(ops.fp_fan_curve(img, a, p, vol, proj), sino)


def test_op_gradients_are_matched_adjoints(case):
    p, a, plan, img, vol, proj = case
    al, pl = a.tolist(), p.tolist()
    w = torch.randn(3, 72, 96, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
    x = img.clone().requires_grad_(True)
    (torch.ops.fbct.fp_fan_curve(x, al, pl, vol, proj) * w).sum().backward()
    assert torch.equal(x.grad, plan.fp_adjoint(w))
    s = plan.fp(img).requires_grad_(True)
    u = torch.randn(3, 64, 64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    (torch.ops.fbct.fbp_fan_curve(s, al, pl, vol, proj) * u).sum().backward()
    assert torch.equal(s.grad, plan.fbp_adjoint(u))
    s2 = plan.fp(img).requires_grad_(True)
    (torch.ops.fbct.bp_fan_curve(s2, al, pl, vol, proj) * u).sum().backward()
    assert torch.equal(s2.grad, plan.bp_adjoint(u))


@pytest.mark.parametrize("name", ["fp_fan_curve", "fp_fan_curve_adjoint", "bp_fan_curve", "fbp_fan_curve",
                                  "back_projection"])
def test_opcheck(case, name):
    p, a, plan, img, vol, proj = case
    x = img if name == "fp_fan_curve" else plan.fp(img)
    x = x.clone().requires_grad_(True)
    torch.library.opcheck(getattr(torch.ops.fbct, name).default, (x, a.tolist(), p.tolist(), vol, proj),
                          test_utils=("test_schema", "test_autograd_registration", "test_faketensor"))
