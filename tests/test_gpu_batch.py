"""GPU: uneven batches (slices of different geometries in one persistent launch).

Checks, through the C ABI (fbct_batch_*):
  * every slice of a batch equals the single-plan FP / FP^T of that slice: bit for bit
    for the launch shape that runs the same kernel body (64-ray bundles, 16x16 tiles),
    within 1e-6 relative L2 for the other shapes (the per-sample arithmetic is the
    same; only the interior/border code path of a pixel can change with the tile);
  * each slice against the FP64 oracle at 1e-5 relative L2 (FP and FP^T);
  * BF16 storage equals the single-plan BF16 path bit for bit;
  * the HIP-graph-captured FP -> residual -> FP^T step equals the eager step on every
    replay, with eager work and new input data between replays;
  * staging the cell tiles' rays in LDS (4 or 8 views per batch) changes no bit of FP^T;
  * the autotuner sets a valid shape and reads it back from its on-disk cache.
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from fanbeamctrecon_tensorflowop_amd import workloads
from fanbeamctrecon_tensorflowop_amd.batch import ADJ_TILE_X, FP_RAYS, StepGraph, UnevenBatch, autotune
from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _small_uneven(n=7, seed=11):
    """Uneven slices small enough for the oracle: sides 32..96, 30..90 views, 40..120 bins."""
    scans = workloads.uneven_scans(n, seed=seed, sides=(32, 48, 64, 96), views=(30, 90), bins=(40, 120),
                                   rfov=40.0, sid=120.0, sdd=240.0)
    phantoms = workloads.uneven_phantoms(scans, seed=seed, rfov=40.0)
    return scans, phantoms


@pytest.fixture(scope="module")
def uneven():
    scans, phantoms = _small_uneven()
    plans = [FanBeamPlan(p, a) for p, a in scans]
    batch = UnevenBatch(plans)
    return scans, phantoms, plans, batch


def _per_slice(batch, packed, s, kind):
    return (batch.sino if kind == "sino" else batch.image)(packed, s).detach().cpu().numpy().astype(np.float64)


def test_batch_layout(uneven):
    scans, phantoms, plans, batch = uneven
    io, so = batch.image_offsets, batch.sino_offsets
    assert io[0] == 0 and so[0] == 0
    for s, p in enumerate(plans):
        assert io[s + 1] - io[s] == p.nx * p.ny
        assert so[s + 1] - so[s] == p.n_views * p.n_cols
    assert batch.updates == sum(p.nx * p.ny * p.n_views for p in plans)
    assert batch.fp_samples == sum(p.fp_samples for p in plans)


def test_batch_equals_single_plans(uneven):
    scans, phantoms, plans, batch = uneven
    x = batch.pack_images(phantoms)
    batch.set_shape(fp_rays=64, fp_wgs_per_cu=8, adj_tile_x=16, adj_wgs_per_cu=8)
    y = batch.fp(x)
    g = batch.fp_adjoint(y)
    torch.cuda.synchronize()
    for s, p in enumerate(plans):
        ys = p.fp(batch.image(x, s)[None])[0]
        gs = p.fp_adjoint(ys[None])[0]
        assert torch.equal(batch.sino(y, s), ys), "slice %d fp" % s
        assert torch.equal(batch.image(g, s), gs), "slice %d fp_adjoint" % s
    _say("batch == single plans for %d slices" % len(plans))


@pytest.mark.parametrize("wgs", [3, 0])  # 0: one workgroup per work item
@pytest.mark.parametrize("rays", FP_RAYS)
@pytest.mark.parametrize("tile", ADJ_TILE_X)
def test_batch_shapes_agree(uneven, rays, tile, wgs):
    scans, phantoms, plans, batch = uneven
    x = batch.pack_images(phantoms)
    batch.set_shape(fp_rays=64, fp_wgs_per_cu=8, adj_tile_x=16, adj_wgs_per_cu=8)
    y0 = batch.fp(x)
    g0 = batch.fp_adjoint(y0)
    batch.set_shape(fp_rays=rays, fp_wgs_per_cu=wgs, adj_tile_x=tile, adj_wgs_per_cu=wgs + 2 if wgs else 0)
    y1 = batch.fp(x)
    g1 = batch.fp_adjoint(y0)
    torch.cuda.synchronize()
    ey = rel_l2(y1.double().cpu().numpy(), y0.double().cpu().numpy())
    eg = rel_l2(g1.double().cpu().numpy(), g0.double().cpu().numpy())
    _say("shape rays=%d tile=%d wgs=%d: fp %.2e adj %.2e" % (rays, tile, wgs, ey, eg))
    assert ey <= 1e-6 and eg <= 1e-6
    batch.set_shape(fp_rays=64, fp_wgs_per_cu=8, adj_tile_x=16, adj_wgs_per_cu=8)


@pytest.mark.parametrize("storage", ["f32", "bf16"])
@pytest.mark.parametrize("wgs", [3, 0])
@pytest.mark.parametrize("views", [4, 8])
def test_batch_staged_cells_bitwise(uneven, views, wgs, storage):
    scans, phantoms, plans, batch = uneven
    dt = torch.float32 if storage == "f32" else torch.bfloat16
    x = batch.pack_images(phantoms)
    batch.set_shape(fp_rays=64, fp_wgs_per_cu=8, adj_tile_x=16, adj_wgs_per_cu=8, adj_stage_views=0)
    y = batch.fp(x).to(dt)
    g0 = batch.fp_adjoint(y)
    batch.set_shape(adj_wgs_per_cu=wgs, adj_stage_views=views)
    g1 = batch.fp_adjoint(y)
    torch.cuda.synchronize()
    batch.set_shape(adj_wgs_per_cu=8, adj_stage_views=0)
    _say("staged cells views=%d wgs=%d %s: max |diff| %.3e" % (views, wgs, storage, float((g1 - g0).abs().max())))
    assert torch.equal(g1, g0)


def test_batch_matches_oracle(uneven, oracle_mod):
    scans, phantoms, plans, batch = uneven
    x = batch.pack_images(phantoms)
    y = batch.fp(x)
    g = batch.fp_adjoint(y)
    torch.cuda.synchronize()
    for s, (p, a) in enumerate(scans):
        o = oracle_mod.Oracle(p, a)
        yo = o.fp(phantoms[s][None].astype(np.float64))[0]
        ys = _per_slice(batch, y, s, "sino")
        go = o.fp_adjoint(ys[None])[0]
        gs = _per_slice(batch, g, s, "image")
        e1, e2 = rel_l2(ys, yo), rel_l2(gs, go)
        _say("slice %d (%dx%d, %d views, %d bins): fp %.2e adj %.2e"
             % (s, plans[s].nx, plans[s].ny, plans[s].n_views, plans[s].n_cols, e1, e2))
        assert e1 <= 1e-5 and e2 <= 1e-5


def test_batch_bf16_storage(uneven):
    scans, phantoms, plans, batch = uneven
    batch.set_shape(fp_rays=64, fp_wgs_per_cu=8, adj_tile_x=16, adj_wgs_per_cu=8)
    xb = batch.pack_images(phantoms, dtype=torch.bfloat16)
    y = batch.fp(xb)
    yb = y.to(torch.bfloat16)
    g = batch.fp_adjoint(yb, out_dtype=torch.bfloat16)
    torch.cuda.synchronize()
    for s, p in enumerate(plans):
        assert torch.equal(batch.sino(y, s), p.fp(batch.image(xb, s)[None])[0]), "slice %d" % s
        ref = p.fp_adjoint(batch.sino(yb, s)[None], out_dtype=torch.bfloat16)[0]
        assert torch.equal(batch.image(g, s), ref), "slice %d adjoint" % s


def test_step_graph_replays_match_eager(uneven):
    scans, phantoms, plans, batch = uneven
    batch.set_shape(fp_rays=32, fp_wgs_per_cu=4, adj_tile_x=32, adj_wgs_per_cu=4)
    x = batch.pack_images(phantoms)
    shifted = [np.roll(ph, 2, axis=1) for ph in phantoms]
    target = batch.fp(batch.pack_images(shifted)).clone()
    graph = StepGraph(batch, x, target)
    rng = np.random.default_rng(5)
    for rep in range(4):
        # eager work between replays, new data written into the static input
        noise = torch.from_numpy(rng.normal(0, 1e-3, size=x.numel()).astype(np.float32)).cuda()
        x.copy_(batch.pack_images(phantoms) + noise)
        _ = batch.fp_adjoint(batch.fp(x))
        torch.cuda.synchronize()
        got = graph.replay().clone()
        torch.cuda.synchronize()
        ref = batch.fp_adjoint(batch.fp(x) - target)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), "replay %d differs from the eager step" % rep
    _say("graph replays == eager")
    batch.set_shape(fp_rays=64, fp_wgs_per_cu=8, adj_tile_x=16, adj_wgs_per_cu=8)


def test_autotune_cache(uneven, tmp_path, monkeypatch):
    scans, phantoms, plans, batch = uneven
    monkeypatch.setenv("FBCT_TUNE_CACHE", str(tmp_path / "shapes.json"))
    x = batch.pack_images(phantoms)
    y = batch.fp(x)
    r1 = autotune(batch, x, y, reps=1)
    assert not r1["cached"]
    assert batch.shape == r1["shape"]
    assert r1["shape"]["fp_rays"] in FP_RAYS and r1["shape"]["adj_tile_x"] in ADJ_TILE_X
    batch.set_shape(fp_rays=64, fp_wgs_per_cu=8, adj_tile_x=16, adj_wgs_per_cu=8)
    r2 = autotune(batch, x, y, reps=1)
    assert r2["cached"] and batch.shape == r1["shape"]
    _say("autotuned %s (fp %.3f ms, adj %.3f ms)" % (r1["shape"], r1["fp_ms"], r1["adj_ms"]))
    batch.set_shape(fp_rays=64, fp_wgs_per_cu=8, adj_tile_x=16, adj_wgs_per_cu=8)
