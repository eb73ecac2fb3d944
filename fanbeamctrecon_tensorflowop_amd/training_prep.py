"""Training-data preparation, the counterpart of FDK_generate in the reference's
FDK/FDK_test_training.py:18-132.

For one exam: the five projection stacks (high, m1, m2, energy, mask; each
``[nslice, nview, ncol]``) are rolled along the view axis so that the scan starts at
the view nearest to the negated start angle (:41-42,88-92), the high-energy stack is
scaled by 1/49000 (:101), slices of high, m1 and m2 are reconstructed with the FDK op
(``BackProjection``, :111-113), and the arrays are written as ``.npy`` files with the
reference's names and shapes (:122-132):

  high_prj_<exam>_<k>.npy  [1, nview*ncol, 1]      high_rcn_<exam>_<k>.npy [1, nimg, nimg, 1]
  m1_prj / m2_prj / eff_prj / mask_prj  (same)     m_rcn_<exam>_<k>.npy    [1, nimg, nimg, 2]

plus ``recon_images/recon_<k-1>.raw`` with the high-energy reconstruction.  The scanner
constants are the script's (:20-31,46-50).  As in the script, the angle list adds the
negated start angle, in degrees, to the view angles in radians (:80-82).
"""
from __future__ import annotations

import math
import os
from types import SimpleNamespace

import numpy as np

__all__ = ["scanner", "view_shift", "roll_views", "fdk_parameters", "fdk_angles", "prepare_exam"]

HIGH_SCALE = 49000.0


def scanner() -> SimpleNamespace:
    """FDK/FDK_test_training.py:20-31,46 constants."""
    ncol = 888
    nfan = 55.0059
    return SimpleNamespace(sdd=946.73233346, sod=538.52, ncol=ncol, nfan=nfan, dcol=nfan / ncol, nimg=512, dz=2.5,
                           shift=1.125, dleft=-27.50295, det_bottom=-35.16105, drow=1.0987828125, rfov=250.0)


def view_shift(start_angle_deg: float, nview: int) -> int:
    """Views to roll by: int(round(-start / (360 / nview))) (:40-42)."""
    return int(round(-start_angle_deg / (360.0 / nview)))


def roll_views(proj: np.ndarray, shift: int) -> np.ndarray:
    """concat(proj[:, shift:], proj[:, :shift]) along the view axis (:88-92)."""
    return np.concatenate((proj[:, shift:proj.shape[1], :], proj[:, 0:shift, :]), axis=1)


def fdk_parameters(nview: int, nview_full: int, sc: SimpleNamespace = None) -> np.ndarray:
    """The 16-float attribute of :61-77 (batch size 1, start angle entry 0)."""
    sc = sc or scanner()
    p = np.arange(16, dtype=np.float32)
    p[:] = [sc.sdd, sc.sod, nview, nview * 360 / nview_full, sc.ncol, sc.nimg, sc.nimg, sc.rfov, sc.dcol, sc.shift,
            sc.dleft, 1, 0, sc.det_bottom, sc.drow, sc.dz]
    return p


def fdk_angles(start_angle_deg: float, nview: int) -> np.ndarray:
    """anglePos[i] = -start + i 2 pi / nview, float32 (:79-82)."""
    f0 = -start_angle_deg
    d = 2.0 * math.pi / nview
    a = np.arange(nview, dtype=np.float32)
    for i in range(nview):
        a[i] = f0 + i * d
    return a


def prepare_exam(index, start_angle_deg, nview_full, stacks: dict, savepath: str, raw_dir: str = None,
                 sc: SimpleNamespace = None):
    """Run the pipeline for one exam.  ``stacks`` maps "high", "m1", "m2", "eff", "mask"
    to float32 arrays [nslice, nview, ncol].  Returns the list of slice ids written."""
    import torch

    from . import ops

    sc = sc or scanner()
    nview = nview_full
    shift = view_shift(start_angle_deg, nview)
    rolled = {k: roll_views(np.asarray(v, dtype=np.float32), shift) for k, v in stacks.items()}
    params = fdk_parameters(nview, nview_full, sc)
    angles = fdk_angles(start_angle_deg, nview)
    nslice = rolled["high"].shape[0]
    nimg, ncol = sc.nimg, sc.ncol
    os.makedirs(savepath, exist_ok=True)
    if raw_dir:
        os.makedirs(raw_dir, exist_ok=True)
    dev = torch.device("cuda")

    def recon(p):
        t = torch.from_numpy(np.ascontiguousarray(p, dtype=np.float32)).to(dev)
        return ops.BackProjection(t, angles, params, [1, nimg, nimg], [1, nview, ncol]).cpu().numpy()

    ids = []
    for i in range(nslice):
        a = rolled["high"][i].reshape(1, nview, ncol) / HIGH_SCALE
        b = rolled["m1"][i].reshape(1, nview, ncol)
        c = rolled["m2"][i].reshape(1, nview, ncol)
        d = rolled["eff"][i].reshape(1, nview, ncol)
        e = rolled["mask"][i].reshape(1, nview, ncol)
        ra, rb, rc = recon(a), recon(b), recon(c)
        sid = "%s_%d" % (index, i + 1)
        ids.append(sid)
        for name, arr in (("high", a), ("m1", b), ("m2", c), ("eff", d), ("mask", e)):
            np.save(os.path.join(savepath, "%s_prj_%s.npy" % (name, sid)), arr.reshape(1, nview * ncol, 1))
        np.save(os.path.join(savepath, "high_rcn_%s.npy" % sid), ra.reshape(1, nimg, nimg, 1))
        m = np.concatenate((rb.reshape(1, nimg, nimg, 1), rc.reshape(1, nimg, nimg, 1)), axis=-1)
        np.save(os.path.join(savepath, "m_rcn_%s.npy" % sid), m)
        if raw_dir:
            from .rawio import write_raw
            write_raw(os.path.join(raw_dir, "recon_%d.raw" % i), ra)
    return ids
