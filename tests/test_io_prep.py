"""Raw-file I/O, the test.py smoke driver and the training-data preparation pipeline
(FDK/FDK_test_training.py counterpart).  CPU tests check file formats, the view roll
and the attribute vectors; GPU tests run the drivers end to end."""
import math
import os

import numpy as np
import pytest

from fanbeamctrecon_tensorflowop_amd import rawio, training_prep


def test_raw_round_trip(tmp_path):
    a = np.random.default_rng(0).normal(size=(3, 8, 6)).astype(np.float32)
    p = tmp_path / "x.raw"
    rawio.write_raw(p, a)
    assert os.path.getsize(p) == a.size * 4
    np.testing.assert_array_equal(rawio.read_binary_dataset(p, [8, 6]), a)
    np.testing.assert_array_equal(rawio.read_projections(p, 3, 8, 6), a)
    # C order, no header: the first bytes are a[0, 0, 0], a[0, 0, 1]
    assert np.frombuffer(open(p, "rb").read(8), dtype=np.float32).tolist() == a[0, 0, :2].tolist()


@pytest.mark.parametrize("start,nview", [(56.4605480091944, 1686), (-10.0, 984), (0.0, 360), (179.9, 720)])
def test_view_roll_matches_numpy_roll(start, nview):
    shift = training_prep.view_shift(start, nview)
    assert shift == int(round(-start / (360.0 / nview)))
    proj = np.arange(2 * nview * 3, dtype=np.float32).reshape(2, nview, 3)
    np.testing.assert_array_equal(training_prep.roll_views(proj, shift), np.roll(proj, -shift, axis=1))


def test_fdk_attributes():
    sc = training_prep.scanner()
    p = training_prep.fdk_parameters(1686, 1686)
    assert p.dtype == np.float32 and p.shape == (16,)
    np.testing.assert_array_equal(p, np.array([946.73233346, 538.52, 1686, 360.0, 888, 512, 512, 250, 55.0059 / 888,
                                               1.125, -27.50295, 1, 0, -35.16105, 1.0987828125, 2.5],
                                              dtype=np.float32))
    a = training_prep.fdk_angles(56.46, 1686)
    # the script adds the negated start angle in degrees to radians
    assert a[0] == np.float32(-56.46) and a[1] == np.float32(-56.46 + 2 * math.pi / 1686)
    assert sc.dcol == 55.0059 / 888


@pytest.mark.gpu
def test_smoke_driver_end_to_end(tmp_path):
    import torch
    from fanbeamctrecon_tensorflowop_amd import compat, smoke_driver
    img, sino, bp, fbp = smoke_driver.run(None, str(tmp_path), slices=2, dim=512)
    cfg = compat.default_config()
    for name, ref, shape in (("test_prj.raw", sino, (2, 984, 888)), ("test_img_BP.raw", bp, (2, 512, 512)),
                             ("test_img_FBP.raw", fbp, (2, 512, 512))):
        got = rawio.read_binary_dataset(tmp_path / name, shape[1:])
        assert got.shape == shape
        np.testing.assert_array_equal(got, ref.cpu().numpy())
    # the same ops called directly
    x = torch.from_numpy(img).cuda()
    np.testing.assert_array_equal(compat.ForwardProjection(x, 2, cfg).cpu().numpy(), sino.cpu().numpy())
    # FBP of a phantom's projections is a recognisable image: correlated with the phantom
    r = np.corrcoef(fbp.cpu().numpy().ravel(), img.ravel())[0, 1]
    assert r > 0.9


@pytest.mark.gpu
def test_training_prep_end_to_end(tmp_path):
    import torch
    from fanbeamctrecon_tensorflowop_amd import ops
    nview, nslice = 360, 2
    rng = np.random.default_rng(1)
    stacks = {k: rng.uniform(0, 1, size=(nslice, nview, 888)).astype(np.float32)
              for k in ("high", "m1", "m2", "eff", "mask")}
    ids = training_prep.prepare_exam(7, 33.3, nview, stacks, str(tmp_path / "npy"), str(tmp_path / "raw"))
    assert ids == ["7_1", "7_2"]
    shift = training_prep.view_shift(33.3, nview)
    p = training_prep.fdk_parameters(nview, nview)
    a = training_prep.fdk_angles(33.3, nview)
    for i, sid in enumerate(ids):
        hp = np.load(tmp_path / "npy" / ("high_prj_%s.npy" % sid))
        assert hp.shape == (1, nview * 888, 1)
        high = np.roll(stacks["high"][i], -shift, axis=0) / 49000
        np.testing.assert_array_equal(hp.reshape(nview, 888), high)
        rcn = np.load(tmp_path / "npy" / ("high_rcn_%s.npy" % sid))
        assert rcn.shape == (1, 512, 512, 1)
        ref = ops.BackProjection(torch.from_numpy(high[None].astype(np.float32)).cuda(), a, p, [1, 512, 512],
                                 [1, nview, 888]).cpu().numpy()
        np.testing.assert_array_equal(rcn.reshape(512, 512), ref.reshape(512, 512))
        m = np.load(tmp_path / "npy" / ("m_rcn_%s.npy" % sid))
        assert m.shape == (1, 512, 512, 2)
        for name in ("m1", "m2", "eff", "mask"):
            assert np.load(tmp_path / "npy" / ("%s_prj_%s.npy" % (name, sid))).shape == (1, nview * 888, 1)
        raw = rawio.read_binary_dataset(tmp_path / "raw" / ("recon_%d.raw" % i), [512, 512])
        np.testing.assert_array_equal(raw[0], rcn.reshape(512, 512))
