"""GPU: the fused FP residual (fbct_fp_residual) equals the projection followed by the
subtraction and the storage conversion, bit for bit: the one-slice, two-slice and
four-slice (two tile classes) FP paths, BF16 and FP32 outputs, FP32 and BF16 images."""
import numpy as np
import pytest
import torch

from fanbeamctrecon_tensorflowop_amd import workloads
from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("geometry,n", [("small", 1), ("small", 2), ("clinical", 4), ("clinical", 5)])
@pytest.mark.parametrize("out_dtype", [torch.bfloat16, torch.float32])
@pytest.mark.parametrize("in_dtype", [torch.float32, torch.bfloat16])
def test_fp_residual_equals_fp_minus_target(geometry, n, out_dtype, in_dtype):
    if geometry == "small":
        p, a = workloads.small_scan()
        img = workloads.ellipse_phantom(256, 256, 128.0, 8, seed=0)
    else:
        p, a = workloads.clinical_scan()
        img = workloads.clinical_phantom(1)[0]
    plan = FanBeamPlan(p, a)
    x = torch.from_numpy(np.stack([np.roll(img, 7 * s, axis=0) for s in range(n)]).astype(np.float32)).cuda()
    x = x.to(in_dtype)
    target = plan.fp(torch.roll(x, shifts=5, dims=2))
    fused = plan.fp_residual(x, target, out_dtype=out_dtype)
    ref = (plan.fp(x) - target).to(out_dtype)
    torch.cuda.synchronize()
    print("fp_residual %s n=%d %s -> %s: max |diff| %.3e" % (geometry, n, in_dtype, out_dtype,
                                                          float((fused.float() - ref.float()).abs().max())), flush=True)
    assert torch.equal(fused, ref)
