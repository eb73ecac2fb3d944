"""End-to-end driver, the counterpart of the reference's test.py (test.py:17-42).

    python -m fanbeamctrecon_tensorflowop_amd.smoke_driver [--input test.fbp] [--outdir .]
                                                           [--slices 2] [--dim 512]

Reads a stack of float32 512x512 slices (C order) -- or, since the reference's input
``test.fbp`` is not distributed, synthesises seeded ellipse phantoms of the same shape
-- and runs ForwardProjection -> BackProjection -> FilteredBackProjection with the
reference's default scan settings (config.py:13-36 through :mod:`compat`), writing
``test_prj.raw``, ``test_img_BP.raw`` and ``test_img_FBP.raw`` as the reference does.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np


def run(input_path=None, outdir=".", slices=2, dim=512, cfg=None):
    import torch

    from . import compat, rawio, workloads

    cfg = cfg or compat.default_config()
    if input_path:
        img = rawio.read_binary_dataset(input_path, [dim, dim])
    else:
        img = np.stack([workloads.ellipse_phantom(dim, dim, cfg.FOVRadius, 8, seed=s) for s in range(slices)])
    nimg = img.shape[0]
    dev = torch.device("cuda")
    x = torch.from_numpy(np.ascontiguousarray(img, dtype=np.float32)).to(dev)
    sino = compat.ForwardProjection(x, nimg, cfg)
    rawio.write_raw(os.path.join(outdir, "test_prj.raw"), sino.cpu().numpy())
    bp = compat.BackProjection(sino, nimg, cfg)
    rawio.write_raw(os.path.join(outdir, "test_img_BP.raw"), bp.cpu().numpy())
    fbp = compat.FilteredBackProjection(sino, nimg, cfg)
    rawio.write_raw(os.path.join(outdir, "test_img_FBP.raw"), fbp.cpu().numpy())
    return img, sino, bp, fbp


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--input", default=None, help="float32 raw stack of dim x dim slices (default: phantoms)")
    ap.add_argument("--outdir", default=".")
    ap.add_argument("--slices", type=int, default=2, help="phantom slices when no input is given")
    ap.add_argument("--dim", type=int, default=512)
    a = ap.parse_args(argv)
    img, sino, bp, fbp = run(a.input, a.outdir, a.slices, a.dim)
    print("smoke_driver: %d slices, sinogram %s, BP %s, FBP %s written to %s"
          % (img.shape[0], tuple(sino.shape), tuple(bp.shape), tuple(fbp.shape), os.path.abspath(a.outdir)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
