"""Benchmark: batched FP + matched-adjoint step on clinical-size fan-beam slices.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--slices 64]

Workload: 512x512 fp32 images (0.7 mm pixels), SID 541 mm, SDD 949 mm, 888 equiangular
bins (1.0239 mm at the detector), 984 views over 360 deg; seeded ellipse phantoms plus
N(0, 1e-4) noise.  One step = forward projection of a batch of slices followed by the
matched adjoint applied to the residual (the forward/backward of a projection loss).

Prints one JSON line.  value = GUPS of the step = 2 * slices * nX*nY*nViews / t (FP and
its adjoint each count nX*nY*nViews voxel-view updates per slice).  Also reported:
FP / BP / FBP GUPS on the batch, ms/slice batched vs the plain per-slice drop-in
calls, the dominant kernel's roofline figures and the FP64 CPU oracle timed on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "GUPS for FP & BP at 512²×984v; batched fwd+adjoint ms/slice vs plain"
# LDS read bandwidth of one MI355X: 96.7 TB/s, measured with a ds_read_b128 throughput
# probe (377 B/ns per CU over 256 CUs); AMD publishes no LDS figure for this part.
LDS_PEAK_GBS = 96700.0
HBM_PEAK_GBS = 8000.0
# Per-kernel HBM traffic (rocprofv3 FETCH_SIZE and WRITE_SIZE passes, calibrated on a plain
# copy) and VALU issue rate (SQ counters) of the clinical 64-slice dispatches; used when the
# workload matches.
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_clinical_64slices.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--slices", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("clinical", "uneven", "fdk"), default="clinical",
                    help="clinical: 64 slices of 512^2 x 984 views (the headline); uneven: 256 slices of "
                         "seeded mixed geometry (sides 256-768, 360-1200 views, 512-1024 bins) in one "
                         "persistent launch per operator, autotuned; fdk: cone-beam FDK of a 512x512x256 "
                         "volume from 720 views of a 1024x256 flat panel, BF16 projections")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="dispatch each operator once on the batch (plus a 256 MiB calibration kernel) "
                         "for counter collection under rocprofv3, then exit")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (slices shard across GPUs "
              "with no communication; see the sharding notes in DESIGN.md). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from fanbeamctrecon_tensorflowop_amd import workloads
    from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.workload == "uneven":
        return uneven_main(args)
    if args.workload == "fdk":
        return fdk_main(args)
    dev = torch.device("cuda:0")
    S = args.slices
    p, a = workloads.clinical_scan(n_slices=S)
    plan = FanBeamPlan(p, a)
    nX, nY, nV, nC = plan.nx, plan.ny, plan.n_views, plan.n_cols
    U = nX * nY * nV                      # voxel-view updates per slice per operator
    samples = plan.fp_samples             # FP samples per slice

    log("bench: building %d phantoms" % S)
    base = workloads.clinical_phantom(min(S, 8))
    img = np.concatenate([np.roll(base, 7 * k, axis=2) for k in range((S + 7) // 8)])[:S]
    x = torch.from_numpy(np.ascontiguousarray(img)).to(dev)
    # target sinograms: projections of the phantoms shifted by 9 pixels (a residual of
    # the size of the signal, as between a prediction and a measurement)
    target = plan.fp(torch.roll(x, shifts=9, dims=2))
    torch.cuda.synchronize()

    if args.profile_kernels:
        # inputs first, then the calibration: an elementwise kernel that reads 256 MiB and
        # writes 256 MiB (262144 KiB each way), for scaling the FETCH_SIZE / WRITE_SIZE
        # counters; after it, each operator runs exactly once
        sino = target.contiguous()
        xb16 = x.to(torch.bfloat16)
        rb16 = (target - plan.fp(xb16)).to(torch.bfloat16)
        src = torch.ones(64 * 2 ** 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.5, out=dst)
        torch.cuda.synchronize()
        for name, fn in (("fp", lambda: plan.fp(x)), ("fp_adjoint", lambda: plan.fp_adjoint(sino)),
                         ("fp_bf16", lambda: plan.fp(xb16)), ("fp_adjoint_bf16", lambda: plan.fp_adjoint(rb16)),
                         ("bp", lambda: plan.bp(sino)), ("fbp", lambda: plan.fbp(sino))):
            fn()
            torch.cuda.synchronize()
            log("profile: %s dispatched" % name)
        print(json.dumps({"profiled": ["fill+copy 256MiB", "fp", "fp_adjoint", "fp_bf16", "fp_adjoint_bf16", "bp",
                                       "fbp"], "slices": S,
                          "fp_samples_per_slice": samples}), flush=True)
        return

    # The step: FP of the batch, residual against a target, matched adjoint of the
    # residual.  BF16 storage: images and the residual sinogram are held in BF16 and
    # widened on load (FP32 arithmetic, FP32 sinogram and gradient outputs).
    xb = x.to(torch.bfloat16)
    targetb = target

    # The residual FP(x) - target is formed in the FP's store (fbct_fp_residual), bitwise
    # equal to projecting, subtracting and converting in three passes.
    def step_f32():
        return plan.fp_adjoint(plan.fp_residual(x, target, out_dtype=torch.float32))

    def step_bf16():
        return plan.fp_adjoint(plan.fp_residual(xb, targetb, out_dtype=torch.bfloat16))

    def timed(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record()
        for _ in range(args.steps):
            out_ = step()
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / args.steps, time.perf_counter() - t0, out_

    log("bench: warmup %d, timing %d steps (BF16 storage)" % (args.warmup, args.steps))
    ms_step, wall, g_bf16 = timed(step_bf16)
    log("bench: FP32 storage step")
    ms_step_f32, _, g_f32 = timed(step_f32)
    gups = 2.0 * S * U / (ms_step * 1e-3) / 1e9

    def time_op(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    reps = max(3, args.steps // 2)
    log("bench: per-operator timings")
    sino = target.contiguous()
    rb = (plan.fp(xb) - targetb).to(torch.bfloat16)
    ms_fp_bf16 = time_op(lambda: plan.fp(xb), reps)  # the step's two launches
    ms_fpa_bf16 = time_op(lambda: plan.fp_adjoint(rb), reps)
    ms_fp = time_op(lambda: plan.fp(x), reps)
    ms_fpa = time_op(lambda: plan.fp_adjoint(sino), reps)
    ms_bp = time_op(lambda: plan.bp(sino), reps)
    ms_fbp = time_op(lambda: plan.fbp(sino), reps)

    log("bench: plain per-slice drop-in")
    xs = [x[s:s + 1].contiguous() for s in range(S)]
    ts = [target[s:s + 1].contiguous() for s in range(S)]

    def plain():
        for s in range(S):
            y = plan.fp(xs[s])
            plan.fp_adjoint(y - ts[s])

    ms_plain = time_op(plain, max(2, reps // 2))

    # Rooflines of the step's two launches.  Algorithmic work of FP and of its adjoint: 4
    # gathered taps of s = 2 B (BF16 storage) per sample and slice; bound: on-chip gathers
    # (LDS, measured peak).  Compulsory HBM bytes: the BF16 input and the FP32 output.
    tr_k = {}
    try:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        if tr.get("slices") == S and tr.get("workload") == "clinical":
            tr_k = tr["kernels"]
    except (OSError, ValueError, KeyError):
        tr_k = {}

    def roofline(kernel, key, ms, hbm_bytes):
        alg = 4 * 2 * samples * S
        ach = alg / (ms * 1e-3) / 1e9
        k = tr_k.get(key, {})
        return {
            "kernel": kernel, "bound": "lds", "achieved": round(ach, 1), "peak": LDS_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / LDS_PEAK_GBS, 4), "traffic": k.get("hbm_bytes"), "launch_ms": round(ms, 4),
            "algorithmic_bytes_per_launch": alg, "compulsory_hbm_bytes_per_launch": hbm_bytes,
            "hbm_frac_compulsory": round(hbm_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 6),
            "valu_instr_per_cu_cycle": k.get("valu_per_cu_cycle"),
            "peak_note": "LDS ds_read_b128 bandwidth measured on MI355X (no published figure); "
                         "valu_instr_per_cu_cycle = SQ_INSTS_VALU / (256 CUs x cycles) from the counter profile",
        }

    rl_adj = roofline("fp_adjoint_cell_staged_kernel<16, bf16, float>", "fp_adjoint_kernel_bf16", ms_fpa_bf16,
                      (2 * nV * nC + 4 * nX * nY) * S)
    rl_fp = roofline("fp_kernel<4, 8, bf16, float>", "fp_kernel_bf16", ms_fp_bf16, (2 * nX * nY + 4 * nV * nC) * S)
    # the dominant kernel of the step (the longer of its two launches) is reported as "roofline"
    rl_adj["model_note"] = ("the byte model counts 4 taps per sample, as a per-sample gather reads them; the "
                            "cell-driven kernel reads a ray's values once per texel cell (6-8 samples), so its "
                            "frac can exceed 1")
    rl_main = rl_fp if ms_fp_bf16 >= ms_fpa_bf16 else rl_adj
    out = {
        "metric": METRIC,
        "value": round(gups, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "clinical fan-beam FP + matched adjoint, %d slices of 512x512, 984 views x 888 bins" % S,
            "slices": S, "image": [nY, nX], "views": nV, "bins": nC,
            "storage": "bf16 image and residual, f32 sinogram and gradient; f32 arithmetic",
        },
        "ms_per_slice_batched_f32_storage": round(ms_step_f32 / S, 4),
        "gups_batched_f32_storage": round(2.0 * S * U / (ms_step_f32 * 1e-3) / 1e9, 3),
        "fp_gups": round(S * U / (ms_fp * 1e-3) / 1e9, 3),
        "fp_bf16_gups": round(S * U / (ms_fp_bf16 * 1e-3) / 1e9, 3),
        "fp_adjoint_gups": round(S * U / (ms_fpa * 1e-3) / 1e9, 3),
        "fp_adjoint_bf16_gups": round(S * U / (ms_fpa_bf16 * 1e-3) / 1e9, 3),
        "bp_gups": round(S * U / (ms_bp * 1e-3) / 1e9, 3),
        "fbp_gups": round(S * U / (ms_fbp * 1e-3) / 1e9, 3),
        "ms_per_slice_batched": round(ms_step / S, 4),
        "ms_per_slice_plain": round(ms_plain / S, 4),
        "batched_vs_plain": round(ms_plain / ms_step, 3),
        "batched_f32_storage_vs_plain": round(ms_plain / ms_step_f32, 3),
        "wall_s": round(wall, 3),
        "roofline": rl_main,
        "roofline_fp": rl_fp,
        "roofline_fp_adjoint": rl_adj,
    }

    if not args.no_cpu_baseline:
        log("bench: CPU oracle baseline (1 slice FP + adjoint)")
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as orc
        o = orc.Oracle(workloads.clinical_scan()[0], a)
        nthr = orc.default_threads()
        one = img[:1]
        t0 = time.perf_counter()
        yo = o.fp(one, nthreads=nthr)
        go = o.fp_adjoint(yo - target[:1].cpu().numpy(), nthreads=nthr)
        tc = time.perf_counter() - t0
        one_b = xb[:1].float().cpu().numpy()  # the BF16-stored image, widened
        yb = o.fp(one_b, nthreads=nthr)
        gb = o.fp_adjoint((yb - target[:1].cpu().numpy()).astype(np.float32).astype(np.float64), nthreads=nthr)
        import numpy as _np

        def _rel(a_, b_):
            return float(_np.linalg.norm(a_ - b_) / _np.linalg.norm(b_))

        out["error_vs_fp64_oracle"] = {
            "note": "relative L2 of slice 0 against the FP64 oracle",
            "step_f32_storage": _rel(g_f32[:1].double().cpu().numpy(), go),
            "step_bf16_storage_vs_fp32_inputs": _rel(g_bf16[:1].double().cpu().numpy(), go),
            "step_bf16_storage_vs_bf16_inputs": _rel(g_bf16[:1].double().cpu().numpy(), gb),
            "fp_f32": _rel(plan.fp(x[:1]).double().cpu().numpy(), yo),
        }
        out["cpu_baseline"] = {
            "value": round(2.0 * U / tc / 1e9, 4), "unit": "GUPS", "cores": nthr, "kind": "port",
            "sample": "FP64 oracle FP + matched adjoint of 1 of the %d slices (%.1f s)" % (S, tc),
        }
    print(json.dumps(out), flush=True)


def uneven_main(args):
    """Uneven batch: 256 slices, each with its own seeded geometry, FP + matched adjoint
    of all slices as one persistent launch per operator (launch shape autotuned)."""
    import numpy as np
    import torch

    from fanbeamctrecon_tensorflowop_amd import workloads
    from fanbeamctrecon_tensorflowop_amd.batch import UnevenBatch, autotune
    from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan

    n = 256
    dev = torch.device("cuda:0")
    log("bench uneven: %d plans" % n)
    scans = workloads.uneven_scans(n, seed=4)
    plans = [FanBeamPlan(p, a) for p, a in scans]
    batch = UnevenBatch(plans)
    log("bench uneven: phantoms")
    phantoms = workloads.uneven_phantoms(scans, seed=4)
    x = batch.pack_images(phantoms, device=dev)
    shifted = [np.roll(ph, max(1, ph.shape[1] // 56), axis=1) for ph in phantoms]
    target = batch.fp(batch.pack_images(shifted, device=dev))
    xb = x.to(torch.bfloat16)
    torch.cuda.synchronize()
    log("bench uneven: autotune")
    tune = autotune(batch, x, target, reps=2)
    log("bench uneven: shape %s" % tune["shape"])

    def step_f32():
        return batch.fp_adjoint(batch.fp(x) - target)

    def step_bf16():
        return batch.fp_adjoint((batch.fp(xb) - target).to(torch.bfloat16))

    def timed(fn, reps):
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms_f32 = timed(step_f32, args.steps)
    ms_bf16 = timed(step_bf16, args.steps)
    ms_fp = timed(lambda: batch.fp(x), max(3, args.steps // 2))
    ms_adj = timed(lambda: batch.fp_adjoint(target), max(3, args.steps // 2))
    log("bench uneven: per-slice plain drop-in")
    xs = [batch.image(x, s)[None] for s in range(n)]
    ts = [batch.sino(target, s)[None] for s in range(n)]

    def plain():
        for s in range(n):
            plans[s].fp_adjoint(plans[s].fp(xs[s]) - ts[s])

    ms_plain = timed(plain, 2)
    U = batch.updates
    out = {
        "metric": "uneven batch FP + matched adjoint: GUPS and ms/slice",
        "value": round(2.0 * U / (ms_bf16 * 1e-3) / 1e9, 3),
        "unit": "GUPS",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_bf16, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "uneven batch of %d slices: side {256,384,512,768}, views 360-1200, bins "
                               "512-1024 (seed 4), FOV 179.2 mm, fan 54.894 deg" % n,
                   "storage": "bf16 image and residual; f32 arithmetic", "slices": n,
                   "voxel_view_updates_per_operator": U, "fp_samples": batch.fp_samples},
        "shape": tune["shape"],
        "ms_per_slice_bf16": round(ms_bf16 / n, 4),
        "ms_per_slice_f32": round(ms_f32 / n, 4),
        "gups_f32": round(2.0 * U / (ms_f32 * 1e-3) / 1e9, 3),
        "fp_gups": round(U / (ms_fp * 1e-3) / 1e9, 3),
        "fp_adjoint_gups": round(U / (ms_adj * 1e-3) / 1e9, 3),
        "ms_per_slice_plain_f32": round(ms_plain / n, 4),
        "batched_bf16_vs_plain": round(ms_plain / ms_bf16, 3),
        "batched_f32_vs_plain": round(ms_plain / ms_f32, 3),
    }
    print(json.dumps(out), flush=True)


def fdk_main(args):
    """Cone-beam FDK: 512x512x256 volume of 0.7 mm voxels, 720 views over 360 deg, flat
    1024x256 panel at 1.0 mm, SID 541 mm, SDD 949 mm; projections are exact line
    integrals of 10 seeded ellipsoids, stored in BF16.  One step = FDK (MFMA row filter
    with cosine pre-weight + voxel-driven backprojection) of the volume."""
    import numpy as np
    import torch

    from fanbeamctrecon_tensorflowop_amd import workloads
    from fanbeamctrecon_tensorflowop_amd.cone import ConePlan

    dev = torch.device("cuda:0")
    g, a = workloads.cone_scan()
    plan = ConePlan(g, a)
    log("bench fdk: projections")
    ells = workloads.random_ellipsoids(10, seed=5)
    proj = workloads.ellipsoid_projections(ells, g, a, device=dev).float()
    proj_b = proj.to(torch.bfloat16)
    torch.cuda.synchronize()
    ws = torch.empty(plan.proj_shape[0] * plan.proj_shape[1] * plan.proj_shape[2] * 4, dtype=torch.uint8,
                     device=dev)

    def timed(fn, reps):
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    log("bench fdk: timing")
    ms = timed(lambda: plan.fdk(proj_b[None], workspace=ws), args.steps)
    ms_f32 = timed(lambda: plan.fdk(proj[None], workspace=ws), args.steps)
    q = plan.filter(proj_b[None])
    ms_filter = timed(lambda: plan.filter(proj_b[None]), args.steps)
    ms_bp = timed(lambda: plan.backproject(q), args.steps)
    vol_b = plan.fdk(proj_b[None], workspace=ws)[0]
    vol_f = plan.fdk(proj[None], workspace=ws)[0]
    U = plan.updates
    rows = g["nViews"] * g["nRows"]
    flops = 2.0 * rows * g["nCols"] ** 2
    out = {
        "metric": "cone-beam FDK GUPS (voxel-view updates/s), 512x512x256 from 720 views",
        "value": round(U / (ms * 1e-3) / 1e9, 3), "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "FDK 512x512x256 (0.7 mm), 720 views, flat 1024x256 @ 1.0 mm, SID 541, SDD 949, "
                               "10 ellipsoids (seed 5), BF16 projections", "voxel_view_updates": U},
        "ms_fdk_f32_projections": round(ms_f32, 3),
        "ms_filter": round(ms_filter, 3),
        "ms_backproject": round(ms - ms_filter, 3),
        "backproject_gups": round(U / ((ms - ms_filter) * 1e-3) / 1e9, 3),
        "ms_backproject_alone": round(ms_bp, 3),
        # dominant kernel: the backprojection; algorithmic work 4 taps x 4 B (FP32 filtered
        # projections) per voxel-view update, read from the LDS-staged detector footprint
        "roofline": {"kernel": "cone_bp_lds_kernel<flat, 32-voxel z-chunks, 2 views per batch, float>", "bound": "lds",
                     "achieved": round(16.0 * U / (ms_bp * 1e-3) / 1e9, 1), "peak": LDS_PEAK_GBS, "unit": "GB/s",
                     "frac": round(16.0 * U / (ms_bp * 1e-3) / 1e9 / LDS_PEAK_GBS, 4), "traffic": None,
                     "launch_ms": round(ms_bp, 4), "algorithmic_bytes_per_launch": 16 * U},
        "filter_roofline": {"bound": "mfma", "achieved": round(flops / (ms_filter * 1e-3) / 1e12, 3),
                            "peak": 2500.0, "unit": "TFLOP/s",
                            "frac": round(flops / (ms_filter * 1e-3) / 1e12 / 2500.0, 5),
                            "note": "nominal 2 rows nCols^2 flops; the six split products run 6x that on MFMA"},
        "bf16_vs_f32_projections_rel_l2": float(torch.linalg.vector_norm((vol_b - vol_f).double()) /
                                                torch.linalg.vector_norm(vol_f.double())),
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as orc
        log("bench fdk: FP64 oracle on sampled lines and slices")
        # filter: 48 (view, row) lines through the oracle's taps and pre-weights
        o = orc.ConeOracle(g, a)
        h, w = o.taps(), o.preweight()
        rng = np.random.default_rng(0)
        pb = proj_b.float().cpu().numpy().astype(np.float64)
        qh = q[0].double().cpu().numpy()
        nC = g["nCols"]
        num = den = 0.0
        for _ in range(48):
            v, r = int(rng.integers(0, g["nViews"])), int(rng.integers(0, g["nRows"]))
            ref = np.convolve(pb[v, r] * w[r], h)[nC - 1:2 * nC - 1]
            num += float(((qh[v, r] - ref) ** 2).sum())
            den += float((ref ** 2).sum())
        err_filter = (num / den) ** 0.5
        # backprojection of the GPU-filtered projections on 3 slices, timed as the CPU baseline
        err_bp = []
        nthr = orc.default_threads()
        t0 = time.perf_counter()
        for k in (40, 128, 200):
            gk = dict(g, nZ=1, zmin=float(np.float32(g["zmin"]) + np.float32(k) * np.float32(g["dz"])))
            ok = orc.ConeOracle(gk, a)
            ref = ok.backproject(qh[None], nthreads=nthr)[0, 0]
            got = vol_b[k].double().cpu().numpy()
            err_bp.append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
        tc = time.perf_counter() - t0
        out["error_vs_fp64_oracle"] = {"filter_48_lines": err_filter, "backproject_slices_40_128_200": err_bp}
        out["cpu_baseline"] = {"value": round(3 * g["nX"] * g["nY"] * g["nViews"] / tc / 1e9, 4), "unit": "GUPS",
                               "cores": nthr, "kind": "port",
                               "sample": "FP64 oracle backprojection of 3 of the 256 slices (%.1f s)" % tc}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
