"""Benchmark: batched FP + matched-adjoint step on clinical-size fan-beam slices, with the
uneven batch and cone-beam FDK workloads beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--slices 64]
                    [--workload all|clinical|uneven|fdk] [--no-cpu-baseline] [--no-counters]

Headline workload: 512x512 fp32 images (0.7 mm pixels), SID 541 mm, SDD 949 mm, 888
equiangular bins (1.0239 mm at the detector), 984 views over 360 deg; seeded ellipse
phantoms plus N(0, 1e-4) noise.  One step = forward projection of 64 slices followed by
the matched adjoint applied to the residual (the forward/backward of a projection loss),
images and residual stored in BF16, FP32 arithmetic.

Prints one JSON line.  value = GUPS of the step = 2 * slices * nX*nY*nViews / t (FP and
its adjoint each count nX*nY*nViews voxel-view updates per slice).  Also reported: FP,
BP, FBP and their transposes' GUPS on the batch, ms/slice batched vs the plain per-slice
drop-in calls, the error of each output against the FP64 oracle, the oracle timed on 4
slices (all host threads) and on 1 slice (one thread), and under "uneven" and "fdk" the
same figures for the 256-slice uneven batch and the 512x512x256 cone-beam FDK.

Rooflines: the event-timed launches of each workload's dominant kernel against the
resource that binds it.  With counters on (the default when rocprofv3 is present), the
bench first runs itself under rocprofv3 in child processes (two counter passes over one
dispatch of every kernel, plus calibration probes whose binding resource is known) and
derives each kernel's HBM traffic, L2 hit rate, VALU-issue and LDS utilisation from
them; those children run before this process touches the GPU.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "GUPS for FP & BP at 512²×984v; batched fwd+adjoint ms/slice vs plain"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
FP32_VALU_PEAK_TFLOPS = 157.3  # MI355X FP32 vector, published
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix, published


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ timing helpers

def timed(torch, fn, reps, warmup=1):
    """Mean ms per call of fn over reps calls, HIP events on the current stream (the
    stream every operator launches on)."""
    for _ in range(max(0, warmup)):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def gups(updates, ms):
    return round(updates / (ms * 1e-3) / 1e9, 3)


# ------------------------------------------------------------------ workload setup

def clinical_setup(torch, S):
    import numpy as np
    from fanbeamctrecon_tensorflowop_amd import workloads
    from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan
    dev = torch.device("cuda:0")
    p, a = workloads.clinical_scan(n_slices=S)
    plan = FanBeamPlan(p, a)
    base = workloads.clinical_phantom(min(S, 8))
    img = np.concatenate([np.roll(base, 7 * k, axis=2) for k in range((S + 7) // 8)])[:S]
    x = torch.from_numpy(np.ascontiguousarray(img)).to(dev)
    # targets: projections of the phantoms shifted by 9 pixels (a residual of the size
    # of the signal, as between a prediction and a measurement)
    target = plan.fp(torch.roll(x, shifts=9, dims=2))
    xb = x.to(torch.bfloat16)
    torch.cuda.synchronize()
    return dict(p=p, a=a, plan=plan, img=img, x=x, xb=xb, target=target)


def uneven_setup(torch, n=256):
    import numpy as np
    from fanbeamctrecon_tensorflowop_amd import workloads
    from fanbeamctrecon_tensorflowop_amd.batch import UnevenBatch
    from fanbeamctrecon_tensorflowop_amd.plan import FanBeamPlan
    dev = torch.device("cuda:0")
    scans = workloads.uneven_scans(n, seed=4)
    plans = [FanBeamPlan(p, a) for p, a in scans]
    batch = UnevenBatch(plans)
    phantoms = workloads.uneven_phantoms(scans, seed=4)
    x = batch.pack_images(phantoms, device=dev)
    shifted = [np.roll(ph, max(1, ph.shape[1] // 56), axis=1) for ph in phantoms]
    target = batch.fp(batch.pack_images(shifted, device=dev))
    xb = x.to(torch.bfloat16)
    torch.cuda.synchronize()
    return dict(scans=scans, plans=plans, batch=batch, phantoms=phantoms, x=x, xb=xb, target=target)


def fdk_setup(torch):
    from fanbeamctrecon_tensorflowop_amd import workloads
    from fanbeamctrecon_tensorflowop_amd.cone import ConePlan
    dev = torch.device("cuda:0")
    g, a = workloads.cone_scan()
    plan = ConePlan(g, a)
    ells = workloads.random_ellipsoids(10, seed=5)
    proj = workloads.ellipsoid_projections(ells, g, a, device=dev).float()
    proj_b = proj.to(torch.bfloat16)
    ws = torch.empty(plan.proj_shape[0] * plan.proj_shape[1] * plan.proj_shape[2] * 4, dtype=torch.uint8,
                     device=dev)
    torch.cuda.synchronize()
    return dict(g=g, a=a, plan=plan, proj=proj, proj_b=proj_b, ws=ws)


# ------------------------------------------------------------ counter profiles
#
# The bench runs itself with --profile-kernels under rocprofv3 (two passes, each within
# the per-block counter limits of one pass).  The child dispatches calibration probes
# and then every operation once, each preceded by a marker kernel whose grid names it.

PROBE_VALU_WGS, PROBE_VALU_ITERS = 8192, 256   # 2^30 v_fma_f32 wave instructions
PROBE_LDS_WGS, PROBE_LDS_ITERS = 8192, 1024    # 2^37 bytes of ds_read_b128
PROBE_COPY_VEC4 = 16 * 2 ** 20                 # 256 MiB read + 256 MiB written

COUNTER_PASSES = (
    ("a", ["FETCH_SIZE", "SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_VALU2",
           "SQ_INSTS_LDS_LOAD_BANDWIDTH", "SQ_LDS_IDX_ACTIVE", "SQ_LDS_BANK_CONFLICT", "SQ_INSTS_VALU_FLOPS_FP32",
           "GRBM_GUI_ACTIVE"]),
    ("b", ["WRITE_SIZE", "TCC_HIT", "TCC_MISS", "SQ_INSTS_LDS_STORE_BANDWIDTH", "SQ_VALU_MFMA_BUSY_CYCLES",
           "SQ_INSTS_MFMA", "SQ_INSTS_VALU_MFMA_MOPS_BF16", "SQ_INSTS_LDS", "SQ_THREAD_CYCLES_VALU",
           "SQ_ACTIVE_INST_VALU", "GRBM_GUI_ACTIVE"]),
)


def _clinical_ops(torch, slices):
    c = clinical_setup(torch, slices)
    plan, x, xb, target = c["plan"], c["x"], c["xb"], c["target"]
    rb = (plan.fp(xb) - target).to(torch.bfloat16)
    sino = target.contiguous()
    return [("clinical.fp_step", lambda: plan.fp_residual(xb, target, out_dtype=torch.bfloat16)),
            ("clinical.fp_adjoint_step", lambda: plan.fp_adjoint(rb)),
            ("clinical.fp", lambda: plan.fp(x)),
            ("clinical.fp_adjoint", lambda: plan.fp_adjoint(sino)),
            ("clinical.bp", lambda: plan.bp(sino)),
            ("clinical.fbp", lambda: plan.fbp(sino)),
            ("clinical.bp_adjoint", lambda: plan.bp_adjoint(x)),
            ("clinical.fbp_adjoint", lambda: plan.fbp_adjoint(x))]


def _uneven_ops(torch):
    u = uneven_setup(torch)
    batch, x, xb, target = u["batch"], u["x"], u["xb"], u["target"]
    rb = (batch.fp(xb) - target).to(torch.bfloat16)
    yo = torch.empty(batch.sino_numel, dtype=torch.float32, device=x.device)
    go = torch.empty(batch.image_numel, dtype=torch.float32, device=x.device)
    return [("uneven.fp_step", lambda: batch.fp(xb, out=yo)),
            ("uneven.fp_adjoint_step", lambda: batch.fp_adjoint(rb, out=go))]


def _fdk_ops(torch):
    f = fdk_setup(torch)
    plan, proj_b = f["plan"], f["proj_b"]
    q = plan.filter(proj_b[None])
    return [("fdk.filter", lambda: plan.filter(proj_b[None])),
            ("fdk.backproject", lambda: plan.backproject(q))]


def profile_ops(torch, workload, slices):
    """[(name, fn)] of the operations a profile dispatches once each (each workload's
    closures in a scope of their own)."""
    ops = []
    if workload in ("all", "clinical"):
        ops += _clinical_ops(torch, slices)
    if workload in ("all", "uneven"):
        ops += _uneven_ops(torch)
    if workload in ("all", "fdk"):
        ops += _fdk_ops(torch)
    return ops


def profile_child(args):
    """Dispatch the probes and each operation once (run under rocprofv3 by the parent)."""
    import ctypes

    import torch
    from fanbeamctrecon_tensorflowop_amd import build as _build
    pl = ctypes.CDLL(_build.PROBE_LIB_PATH)
    ops = profile_ops(torch, args.workload, args.slices)
    dev = torch.device("cuda:0")
    pout = torch.empty(max(PROBE_VALU_WGS, PROBE_LDS_WGS) * 256, dtype=torch.float32, device=dev)
    src = torch.ones(PROBE_COPY_VEC4 * 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    marks = torch.empty(4096, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = ctypes.c_void_p
    names = ["probe.valu", "probe.lds", "probe.copy"] + [n for n, _ in ops]
    fns = [lambda: pl.fbct_probe_valu(P(pout.data_ptr()), PROBE_VALU_WGS, PROBE_VALU_ITERS, st),
           lambda: pl.fbct_probe_lds(P(pout.data_ptr()), PROBE_LDS_WGS, PROBE_LDS_ITERS, st),
           lambda: pl.fbct_probe_copy(P(src.data_ptr()), P(dst.data_ptr()), ctypes.c_longlong(PROBE_COPY_VEC4), st)]
    fns += [fn for _, fn in ops]
    for tag, fn in enumerate(fns):
        pl.fbct_probe_marker(P(marks.data_ptr()), tag, st)
        fn()
        torch.cuda.synchronize()
    print(json.dumps({"profiled": names}), flush=True)


def _read_counter_csv(path):
    """{dispatch_id: {"kernel": name, "grid": n, "counters": {name: value}}}"""
    out = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            d = int(row.get("Dispatch_Id") or row.get("Dispatch_ID") or 0)
            e = out.setdefault(d, {"kernel": row.get("Kernel_Name", ""), "grid": int(float(row.get("Grid_Size", 0))),
                                   "counters": {}, "ns": 0.0})
            try:
                e["ns"] = float(row["End_Timestamp"]) - float(row["Start_Timestamp"])
            except (KeyError, ValueError):
                pass
            name = row.get("Counter_Name", "")
            try:
                val = float(row.get("Counter_Value", "nan"))
            except ValueError:
                continue
            e["counters"][name] = e["counters"].get(name, 0.0) + val
    return out


def _group_by_marker(disp, names):
    """Counter sums per profiled operation (dispatches after a marker belong to it)."""
    ops, cur = {}, None
    for d in sorted(disp):
        e = disp[d]
        if "probe_marker_kernel" in e["kernel"]:
            tag = e["grid"] // 64 - 1
            cur = names[tag] if 0 <= tag < len(names) else None
            continue
        if cur is None:
            continue
        o = ops.setdefault(cur, {"kernels": [], "counters": {}, "ns": 0.0})
        # the name without its argument list ("(anonymous namespace)" holds a parenthesis)
        o["kernels"].append(e["kernel"].replace("(anonymous namespace)::", "").split("(")[0])
        o["ns"] += e["ns"]
        for k, v in e["counters"].items():
            o["counters"][k] = o["counters"].get(k, 0.0) + v
    return ops


def collect_counters(args):
    """Run the counter passes in child processes; {op: {"kernels": [...], "counters": {...}}}
    or {"error": reason}.  Runs before this process touches the GPU."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return {"error": "rocprofv3 not found"}
    tmp = tempfile.mkdtemp(prefix="fbct_prof_")
    merged, names = {}, None
    try:
        for tag, counters in COUNTER_PASSES:
            outdir = os.path.join(tmp, tag)
            cmd = [exe, "--pmc"] + counters + ["-d", outdir, "-o", tag, "--output-format", "csv", "--",
                                               sys.executable, os.path.abspath(__file__), "--profile-kernels",
                                               "--workload", args.workload, "--slices", str(args.slices)]
            log("bench: counter pass %s (%d counters)" % (tag, len(counters)))
            t0 = time.perf_counter()
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
            log("bench: counter pass %s rc=%d in %.0f s" % (tag, r.returncode, time.perf_counter() - t0))
            line = [ln for ln in r.stdout.splitlines() if ln.startswith("{\"profiled\"")]
            files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
            if not line or not files:
                err = [ln for ln in (r.stderr or "").splitlines() if not ln[:1] in ("W", "E", "I") or
                       "Error" in ln or "error" in ln]
                return {"error": "counter pass %s produced no results (rc %d): %s" % (
                    tag, r.returncode, " | ".join(err[-12:])[-1500:])}
            names = json.loads(line[-1])["profiled"]
            ops = _group_by_marker(_read_counter_csv(files[0]), names)
            for op, v in ops.items():
                m = merged.setdefault(op, {"kernels": v["kernels"], "counters": {}, "ns": {}})
                # a counter read again in a later pass keeps its first reading; the later
                # one is kept under name_<pass>
                for k, x in v["counters"].items():
                    m["counters"][k if k not in m["counters"] else "%s_%s" % (k, tag)] = x
                m["ns"][tag] = v["ns"]
        return merged
    except (OSError, subprocess.SubprocessError, ValueError) as e:
        return {"error": "counter passes failed: %s" % e}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


# ------------------------------------------------------------------ the workloads

def bench_clinical(torch, args):
    """The headline: BF16-storage FP + matched adjoint of 64 clinical slices."""
    import numpy as np
    c = clinical_setup(torch, args.slices)
    S, plan, x, xb, target, img = args.slices, c["plan"], c["x"], c["xb"], c["target"], c["img"]
    nX, nY, nV, nC = plan.nx, plan.ny, plan.n_views, plan.n_cols
    U = nX * nY * nV                      # voxel-view updates per slice per operator

    # The step: FP of the batch with the residual against the target formed in the FP's
    # store (fbct_fp_residual, bitwise equal to projecting, subtracting and converting in
    # three passes), then the matched adjoint of the residual.  BF16 storage: the images
    # and the residual are held in BF16 and widened on load (FP32 arithmetic, FP32 gradient).
    def step_bf16():
        return plan.fp_adjoint(plan.fp_residual(xb, target, out_dtype=torch.bfloat16))

    def step_f32():
        return plan.fp_adjoint(plan.fp_residual(x, target, out_dtype=torch.float32))

    log("bench: warmup %d, timing %d steps (BF16 storage)" % (args.warmup, args.steps))
    t0 = time.perf_counter()
    ms_step = timed(torch, step_bf16, args.steps, args.warmup)
    wall = time.perf_counter() - t0
    ms_step_f32 = timed(torch, step_f32, args.steps, args.warmup)
    g_bf16, g_f32 = step_bf16(), step_f32()

    reps = max(3, args.steps // 2)
    log("bench: per-operator timings")
    sino = target.contiguous()
    rb = plan.fp_residual(xb, target, out_dtype=torch.bfloat16)
    ms = {
        "fp_step": timed(torch, lambda: plan.fp_residual(xb, target, out_dtype=torch.bfloat16), reps),
        "fp_adjoint_step": timed(torch, lambda: plan.fp_adjoint(rb), reps),
        "fp": timed(torch, lambda: plan.fp(x), reps),
        "fp_bf16": timed(torch, lambda: plan.fp(xb), reps),
        "fp_adjoint": timed(torch, lambda: plan.fp_adjoint(sino), reps),
        "bp": timed(torch, lambda: plan.bp(sino), reps),
        "fbp": timed(torch, lambda: plan.fbp(sino), reps),
        "bp_adjoint": timed(torch, lambda: plan.bp_adjoint(x), reps),
        "fbp_adjoint": timed(torch, lambda: plan.fbp_adjoint(x), reps),
    }
    log("bench: plain per-slice drop-in")
    xs = [x[s:s + 1].contiguous() for s in range(S)]
    ts = [target[s:s + 1].contiguous() for s in range(S)]

    def plain():
        for s in range(S):
            plan.fp_adjoint(plan.fp(xs[s]) - ts[s])

    ms_plain = timed(torch, plain, max(2, reps // 2))
    out = {
        "metric": METRIC,
        "value": round(2.0 * S * U / (ms_step * 1e-3) / 1e9, 3),
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
            "storage": "bf16 image and residual, f32 gradient; f32 arithmetic",
            "voxel_view_updates_per_operator": S * U, "fp_samples_per_operator": S * plan.fp_samples,
        },
        "ms_per_slice_batched": round(ms_step / S, 4),
        "ms_per_slice_batched_f32_storage": round(ms_step_f32 / S, 4),
        "gups_batched_f32_storage": round(2.0 * S * U / (ms_step_f32 * 1e-3) / 1e9, 3),
        "ms_per_slice_plain": round(ms_plain / S, 4),
        "batched_vs_plain": round(ms_plain / ms_step, 3),
        "batched_f32_storage_vs_plain": round(ms_plain / ms_step_f32, 3),
        "gups": {k: gups(S * U, v) for k, v in ms.items()},
        "ms_per_call": {k: round(v, 4) for k, v in ms.items()},
        "wall_s": round(wall, 3),
    }

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as orc
        o = orc.Oracle(c["p"], c["a"])
        nthr = orc.default_threads()
        k = min(4, S)
        log("bench: CPU oracle baseline (%d slices FP + adjoint, %d threads; 1 slice, 1 thread)" % (k, nthr))
        tgt = target[:k].double().cpu().numpy()
        t0 = time.perf_counter()
        yo = o.fp(img[:k], nthreads=nthr)
        go = o.fp_adjoint(yo - tgt, nthreads=nthr)
        tn = time.perf_counter() - t0
        t0 = time.perf_counter()
        g1 = o.fp_adjoint(o.fp(img[:1], nthreads=1) - tgt[:1], nthreads=1)
        t1 = time.perf_counter() - t0
        assert np.allclose(g1, go[:1], rtol=1e-9, atol=1e-12 * np.abs(go[:1]).max())
        xb_h = xb[:k].float().cpu().numpy()  # the BF16-stored images, widened
        gb = o.fp_adjoint((o.fp(xb_h, nthreads=nthr) - tgt).astype(np.float32).astype(np.float64), nthreads=nthr)
        out["error_vs_fp64_oracle"] = {
            "note": "relative L2 of the first %d slices' outputs against the FP64 oracle" % k,
            "step_f32_storage": rel_l2(g_f32[:k].double().cpu().numpy(), go),
            "step_bf16_storage_vs_fp32_inputs": rel_l2(g_bf16[:k].double().cpu().numpy(), go),
            "step_bf16_storage_vs_bf16_inputs": rel_l2(g_bf16[:k].double().cpu().numpy(), gb),
            "fp_f32": rel_l2(plan.fp(x[:k]).double().cpu().numpy(), yo),
        }
        out["cpu_baseline"] = {
            "value": round(2.0 * k * U / tn / 1e9, 4), "unit": "GUPS", "cores": nthr, "kind": "port",
            "sample": "FP64 oracle FP + matched adjoint of %d of the %d slices on %d threads (%.1f s); scaled "
                      "linearly the %d-slice step takes %.1f s" % (k, S, nthr, tn, S, tn * S / k),
            "ms_per_slice": round(tn / k * 1e3, 1),
            "one_thread": {"value": round(2.0 * U / t1 / 1e9, 4), "unit": "GUPS", "cores": 1,
                           "sample": "1 slice FP + matched adjoint, 1 thread (%.1f s)" % t1,
                           "ms_per_slice": round(t1 * 1e3, 1)},
        }
    return out, dict(S=S, U=U, samples=plan.fp_samples, nX=nX, nY=nY, nV=nV, nC=nC, ms=ms)


def bench_uneven(torch, args):
    """The 256-slice uneven batch: FP + matched adjoint of all slices as one persistent
    launch per operator, launch shape autotuned."""
    import numpy as np
    from fanbeamctrecon_tensorflowop_amd.batch import autotune
    log("bench uneven: setup (256 plans, phantoms)")
    u = uneven_setup(torch)
    batch, plans, x, xb, target = u["batch"], u["plans"], u["x"], u["xb"], u["target"]
    n = len(plans)
    log("bench uneven: autotune")
    tune = autotune(batch, x, target, reps=2)

    def step_bf16():
        return batch.fp_adjoint((batch.fp(xb) - target).to(torch.bfloat16))

    def step_f32():
        return batch.fp_adjoint(batch.fp(x) - target)

    ms_bf16 = timed(torch, step_bf16, args.steps, args.warmup)
    ms_f32 = timed(torch, step_f32, args.steps, args.warmup)
    g_bf16, g_f32 = step_bf16(), step_f32()
    reps = max(3, args.steps // 2)
    rb = (batch.fp(xb) - target).to(torch.bfloat16)
    ms = {"fp_step": timed(torch, lambda: batch.fp(xb), reps),
          "fp_adjoint_step": timed(torch, lambda: batch.fp_adjoint(rb), reps),
          "fp": timed(torch, lambda: batch.fp(x), reps),
          "fp_adjoint": timed(torch, lambda: batch.fp_adjoint(target), reps)}
    log("bench uneven: per-slice plain drop-in")
    xs = [batch.image(x, s)[None] for s in range(n)]
    ts = [batch.sino(target, s)[None] for s in range(n)]

    def plain():
        for s in range(n):
            plans[s].fp_adjoint(plans[s].fp(xs[s]) - ts[s])

    ms_plain = timed(torch, plain, 2)
    U = batch.updates
    out = {
        "metric": "uneven batch FP + matched adjoint: GUPS and ms/slice",
        "value": round(2.0 * U / (ms_bf16 * 1e-3) / 1e9, 3), "unit": "GUPS",
        "ms_per_step": round(ms_bf16, 3), "higher_is_better": True, "dtype": "f32",
        "config": {"workload": "uneven batch of %d slices: side {256,384,512,768}, views 360-1200, bins "
                               "512-1024 (seed 4), FOV 179.2 mm, fan 54.894 deg" % n,
                   "storage": "bf16 image and residual; f32 arithmetic", "slices": n,
                   "voxel_view_updates_per_operator": U, "fp_samples": batch.fp_samples},
        "shape": tune["shape"],
        "ms_per_slice": round(ms_bf16 / n, 4),
        "ms_per_slice_f32_storage": round(ms_f32 / n, 4),
        "gups_f32_storage": round(2.0 * U / (ms_f32 * 1e-3) / 1e9, 3),
        "gups": {k: gups(U, v) for k, v in ms.items()},
        "ms_per_call": {k: round(v, 4) for k, v in ms.items()},
        "ms_per_slice_plain": round(ms_plain / n, 4),
        "batched_vs_plain": round(ms_plain / ms_bf16, 3),
        "batched_f32_storage_vs_plain": round(ms_plain / ms_f32, 3),
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as orc
        nthr = orc.default_threads()
        pick = [0, n // 3, 2 * n // 3, n - 1]
        log("bench uneven: CPU oracle on slices %s" % pick)
        errs, errs_b, upd = [], [], 0
        t0 = time.perf_counter()
        refs = []
        for s in pick:
            p, a = u["scans"][s]
            o = orc.Oracle(p, a)
            tg = batch.sino(target, s).double().cpu().numpy()[None]
            refs.append((o, tg, o.fp_adjoint(o.fp(u["phantoms"][s][None], nthreads=nthr) - tg, nthreads=nthr)))
            upd += plans[s].nx * plans[s].ny * plans[s].n_views
        tn = time.perf_counter() - t0
        for s, (o, tg, go) in zip(pick, refs):
            errs.append(rel_l2(batch.image(g_f32, s).double().cpu().numpy(), go[0]))
            xbs = batch.image(xb, s).float().cpu().numpy()[None]
            gb = o.fp_adjoint((o.fp(xbs, nthreads=nthr) - tg).astype(np.float32).astype(np.float64), nthreads=nthr)
            errs_b.append(rel_l2(batch.image(g_bf16, s).double().cpu().numpy(), gb[0]))
        out["error_vs_fp64_oracle"] = {"note": "relative L2 of the step's gradient for slices %s" % pick,
                                       "step_f32_storage": errs, "step_bf16_storage_vs_bf16_inputs": errs_b}
        out["cpu_baseline"] = {"value": round(2.0 * upd / tn / 1e9, 4), "unit": "GUPS", "cores": nthr,
                               "kind": "port",
                               "sample": "FP64 oracle FP + matched adjoint of slices %s of the %d (%.1f s); "
                                         "scaled by updates the batch step takes %.0f s" % (
                                             pick, n, tn, tn * U / upd)}
    return out, dict(U=U, samples=batch.fp_samples, n=n, ms=ms)


def bench_fdk(torch, args):
    """Cone-beam FDK: 512x512x256 volume of 0.7 mm voxels, 720 views over 360 deg, flat
    1024x256 panel at 1.0 mm, SID 541 mm, SDD 949 mm; projections are exact line integrals
    of 10 seeded ellipsoids, stored in BF16.  One step = FDK (MFMA row filter with cosine
    pre-weight + voxel-driven backprojection) of the volume."""
    import numpy as np
    log("bench fdk: setup")
    f = fdk_setup(torch)
    g, a, plan, proj, proj_b, ws = f["g"], f["a"], f["plan"], f["proj"], f["proj_b"], f["ws"]
    ms_step = timed(torch, lambda: plan.fdk(proj_b[None], workspace=ws), args.steps, args.warmup)
    ms_f32 = timed(torch, lambda: plan.fdk(proj[None], workspace=ws), args.steps, args.warmup)
    q = plan.filter(proj_b[None])
    ms = {"filter": timed(torch, lambda: plan.filter(proj_b[None]), args.steps),
          "backproject": timed(torch, lambda: plan.backproject(q), args.steps)}
    vol_b = plan.fdk(proj_b[None], workspace=ws)[0]
    vol_f = plan.fdk(proj[None], workspace=ws)[0]
    U = plan.updates
    out = {
        "metric": "cone-beam FDK GUPS (voxel-view updates/s), 512x512x256 from 720 views",
        "value": gups(U, ms_step), "unit": "GUPS", "ms_per_step": round(ms_step, 3), "higher_is_better": True,
        "dtype": "f32",
        "config": {"workload": "FDK 512x512x256 (0.7 mm), 720 views, flat 1024x256 @ 1.0 mm, SID 541, SDD 949, "
                               "10 ellipsoids (seed 5), BF16 projections", "voxel_view_updates": U},
        "ms_fdk_f32_projections": round(ms_f32, 3),
        "ms_per_call": {k: round(v, 4) for k, v in ms.items()},
        "backproject_gups": gups(U, ms["backproject"]),
        "bf16_vs_f32_projections_rel_l2": float(torch.linalg.vector_norm((vol_b - vol_f).double()) /
                                                torch.linalg.vector_norm(vol_f.double())),
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as orc
        log("bench fdk: FP64 oracle on sampled lines and slices")
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
        err_bp = []
        nthr = orc.default_threads()
        ks = (40, 100, 160, 220)
        t0 = time.perf_counter()
        for k in ks:
            gk = dict(g, nZ=1, zmin=float(np.float32(g["zmin"]) + np.float32(k) * np.float32(g["dz"])))
            ref = orc.ConeOracle(gk, a).backproject(qh[None], nthreads=nthr)[0, 0]
            err_bp.append(rel_l2(vol_b[k].double().cpu().numpy(), ref))
        tc = time.perf_counter() - t0
        out["error_vs_fp64_oracle"] = {"filter_48_lines": (num / den) ** 0.5, "backproject_slices": list(ks),
                                       "backproject": err_bp}
        upd = len(ks) * g["nX"] * g["nY"] * g["nViews"]
        out["cpu_baseline"] = {"value": round(upd / tc / 1e9, 4), "unit": "GUPS", "cores": nthr, "kind": "port",
                               "sample": "FP64 oracle backprojection of %d of the 256 slices (%.1f s); scaled, the "
                                         "volume takes %.0f s" % (len(ks), tc, tc * U / upd)}
    return out, dict(U=U, ms=ms, rows=g["nViews"] * g["nRows"], nC=g["nCols"])


# --------------------------------------------------------------- rooflines

def _probe_rates(prof):
    out = {}
    try:
        ns = {k: prof[k]["ns"].get("a") for k in ("probe.valu", "probe.lds", "probe.copy")}
        if ns["probe.valu"]:
            out["fp32_fma_tflops"] = round(2.0 * 64 * PROBE_VALU_WGS * 256 / 64 * PROBE_VALU_ITERS * 16 * 8 /
                                           ns["probe.valu"] / 1e3, 1)
        if ns["probe.lds"]:
            out["lds_read_tbs"] = round(16.0 * PROBE_LDS_WGS * 256 * PROBE_LDS_ITERS * 4 / ns["probe.lds"] / 1e3, 1)
        if ns["probe.copy"]:
            out["hbm_copy_gbs"] = round(32.0 * PROBE_COPY_VEC4 / ns["probe.copy"], 1)
    except (KeyError, TypeError, AttributeError):
        pass
    return out


def calibration_summary(cal):
    """The calibration record as printed: scale factors rounded, probe rates as measured."""
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in cal.items()}


def _calibration(prof):
    """Scale factors and probe readings of a counter profile, or None.  GRBM_GUI_ACTIVE
    sums the 8 XCDs' busy cycles, so per-GPU-cycle rates divide by it / 8."""
    try:
        pv, pl, pc = (prof[k]["counters"] for k in ("probe.valu", "probe.lds", "probe.copy"))
        copy_bytes = 16.0 * PROBE_COPY_VEC4
        return {
            "fetch_factor": copy_bytes / (pc["FETCH_SIZE"] * 1024.0),
            "write_factor": copy_bytes / (pc["WRITE_SIZE"] * 1024.0),
            # the FMA probe keeps every SIMD issuing a VALU instruction each quad-cycle
            "probe_valu_busy": _busy(pv)["valu_issue"],
            # the conflict-free LDS probe: its LDS pipe activity and bytes per active cycle
            "probe_lds_busy": _busy(pl)["lds_pipe"],
            "lds_bytes_per_active_cycle": 64.0 * pl["SQ_INSTS_LDS_LOAD_BANDWIDTH"] / pl["SQ_LDS_IDX_ACTIVE"],
            # SQ_INSTS_VALU_FLOPS_FP32 counts per wave instruction: flops = counter x 64
            "flops_per_count": 2.0 * 64.0 * 2 ** 30 / pv["SQ_INSTS_VALU_FLOPS_FP32"] if pv.get(
                "SQ_INSTS_VALU_FLOPS_FP32") else 64.0,
            # the probes' own rates (kernel time from the first pass's trace): non-packed
            # FP32 FMA issue, conflict-free 16-byte LDS reads, a plain HBM copy
            "measured_peaks": _probe_rates(prof),
        }
    except (KeyError, TypeError, ZeroDivisionError):
        return None


def _busy(c):
    """Busy fractions of one dispatch from its counters: VALU issue (quad-cycles in which
    a SIMD issued a VALU instruction, two issued together counting once, over the 1024
    SIMDs' quad-cycles), LDS pipe (SQ_LDS_IDX_ACTIVE, bank-conflict cycles included, over
    the 256 CUs' cycles) and the matrix cores (SQ_VALU_MFMA_BUSY_CYCLES over the SIMDs'
    cycles)."""
    g = c["GRBM_GUI_ACTIVE"]  # 8 x GPU cycles
    b = {"valu_issue": (c["SQ_ACTIVE_INST_VALU"] - c.get("SQ_ACTIVE_INST_VALU2", 0.0)) / g / 32.0,
         "lds_pipe": c["SQ_LDS_IDX_ACTIVE"] / g / 32.0}
    if c.get("SQ_INSTS_MFMA"):
        b["mfma"] = c.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / (c.get("GRBM_GUI_ACTIVE_b", g) / 8.0) / 1024.0
    return b


# LDS: 256 B per active cycle per CU (the width the conflict-free probe moves per cycle
# of SQ_LDS_IDX_ACTIVE), 256 CUs at 2.4 GHz; AMD publishes no LDS figure for MI355X
LDS_PEAK_GBS = 256.0 * 256 * 2.4


def roofline(kernel, ms, op, prof, cal, flops=None, lds_bytes=None, model=""):
    """Roofline entry of one event-timed launch (ms per call) of profiled operation op.

    bound: the busiest resource measured by the counters -- VALU issue, LDS pipe, matrix
    cores, or HBM (FETCH_SIZE and WRITE_SIZE scaled on a plain copy, over the launch
    time, against 8 TB/s).  achieved: the path's algorithmic work per launch (flops for
    VALU and the matrix cores, LDS bytes for the LDS pipe, measured bytes for HBM) over
    the event-timed launch; peak: MI355X's published figure for that resource (FP32
    vector 157.3 TFLOP/s, dense BF16 matrix 2.5 PFLOP/s, HBM 8 TB/s; LDS as above)."""
    r = {"kernel": kernel, "launch_ms": round(ms, 4)}
    sec = ms * 1e-3
    alg = {"model": model}
    if flops:
        alg["flops_per_launch"] = flops
        alg["tflops"] = round(flops / sec / 1e12, 3)
    if lds_bytes:
        alg["lds_bytes_per_launch"] = lds_bytes
        alg["lds_gbs"] = round(lds_bytes / sec / 1e9, 1)
    r["algorithmic"] = alg
    e = prof.get(op) if isinstance(prof, dict) else None
    if not e or not cal:
        r.update({"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None, "traffic": None,
                  "note": "no counter profile: %s" % (prof.get("error") if isinstance(prof, dict) else "disabled")})
        return r
    # the kernels the profile saw for this operation (their launches, in order, once each)
    names = []
    for k in e["kernels"]:
        k = k.replace("void ", "").replace("fbct::", "").replace("unsigned short", "bf16")
        if k not in names:
            names.append(k)
    r["kernel"] = " + ".join(names) if names else kernel
    c = e["counters"]
    traffic = c.get("FETCH_SIZE", 0.0) * 1024.0 * cal["fetch_factor"] + c.get("WRITE_SIZE", 0.0) * 1024.0 * cal[
        "write_factor"]
    busy = _busy(c)
    busy["hbm"] = traffic / sec / 1e9 / HBM_PEAK_GBS
    # the busiest resource; between equal readings the one the path's model prices
    priced = {"valu_issue": bool(flops), "mfma": bool(flops), "lds_pipe": bool(lds_bytes), "hbm": True}
    bound = max(busy, key=lambda k: (round(busy[k], 3), priced.get(k, False)))
    measured_lds = 64.0 * (c.get("SQ_INSTS_LDS_LOAD_BANDWIDTH", 0.0) + c.get("SQ_INSTS_LDS_STORE_BANDWIDTH", 0.0))
    if bound == "valu_issue" and flops:
        ach, peak, unit = flops / sec / 1e12, FP32_VALU_PEAK_TFLOPS, "TFLOP/s"
    elif bound == "mfma" and flops:
        ach, peak, unit = flops / sec / 1e12, BF16_MFMA_PEAK_TFLOPS, "TFLOP/s"
    elif bound == "lds_pipe":
        # the model's LDS bytes, else the bytes the counters saw moving
        ach, peak, unit = (lds_bytes or measured_lds) / sec / 1e9, LDS_PEAK_GBS, "GB/s"
    else:
        ach, peak, unit = traffic / sec / 1e9, HBM_PEAK_GBS, "GB/s"
    hit, miss = c.get("TCC_HIT"), c.get("TCC_MISS")
    r.update({
        "bound": bound, "achieved": round(ach, 3), "peak": peak, "unit": unit, "frac": round(ach / peak, 4),
        "traffic": int(traffic), "hbm_gbs": round(traffic / sec / 1e9, 1),
        "busy": {k: round(v, 4) for k, v in busy.items()},
        "l2_hit_rate": round(hit / (hit + miss), 4) if hit is not None and miss is not None and hit + miss > 0
        else None,
        "lds_bank_conflict_per_active": round(c.get("SQ_LDS_BANK_CONFLICT", 0.0) / c["SQ_LDS_IDX_ACTIVE"], 4)
        if c.get("SQ_LDS_IDX_ACTIVE") else None,
        "valu_dual_issue_per_issue": round(c.get("SQ_ACTIVE_INST_VALU2", 0.0) / c["SQ_ACTIVE_INST_VALU"], 4)
        if c.get("SQ_ACTIVE_INST_VALU") else None,
        # active lanes per VALU instruction (divergence): thread-cycles over wave-cycles x 64,
        # both from the second pass
        "valu_lane_utilisation": round(c["SQ_THREAD_CYCLES_VALU"] / (64.0 * c["SQ_ACTIVE_INST_VALU_b"]), 4)
        if c.get("SQ_THREAD_CYCLES_VALU") and c.get("SQ_ACTIVE_INST_VALU_b") else None,
        "measured": {"lds_gbs": round(measured_lds / sec / 1e9, 1),
                     "fp32_tflops": round(c.get("SQ_INSTS_VALU_FLOPS_FP32", 0.0) * cal["flops_per_count"] / sec / 1e12,
                                          3)},
        "dispatches": len(e["kernels"]),
    })
    return r


# --------------------------------------------------------------------- main

def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--slices", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (rooflines then carry the timings only)")
    ap.add_argument("--workload", choices=("all", "clinical", "uneven", "fdk"), default="all",
                    help="all (default): the clinical headline plus 'uneven' and 'fdk' sub-objects; "
                         "clinical: 64 slices of 512^2 x 984 views; uneven: 256 slices of seeded mixed geometry "
                         "(sides 256-768, 360-1200 views, 512-1024 bins) in one persistent launch per operator, "
                         "autotuned; fdk: cone-beam FDK of a 512x512x256 volume from 720 views of a 1024x256 "
                         "flat panel, BF16 projections")
    ap.add_argument("--counters-out", default=None,
                    help="also write the per-operation counter sums of the rocprofv3 passes (JSON) to this file")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="dispatch the calibration probes and each operator of the workload once (for rocprofv3 "
                         "counter passes), then exit")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (slices shard across GPUs "
              "with no communication). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.profile_kernels:
        return profile_child(args)

    # counter passes first: child processes, before this one touches the GPU
    prof = {"error": "disabled (--no-counters)"} if args.no_counters else collect_counters(args)
    if args.counters_out:
        with open(args.counters_out, "w") as f:
            json.dump(prof, f, indent=1, sort_keys=True)
    cal = _calibration(prof) if "error" not in prof else None

    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    result = {}
    if args.workload in ("all", "clinical"):
        out, ctx = bench_clinical(torch, args)
        S, samples, ms, U = ctx["S"], ctx["samples"], ctx["ms"], ctx["U"]
        # FP and FP^T: a bilinear sample is 4 taps times weights (8 flops) per slice, and
        # reads 4 taps (2 B each in BF16 storage, 4 B in FP32); the cell-walk kernels do
        # less (a texel cell's taps and weights once for its 6-8 samples)
        fpm = "8 flops and 4 taps per FP sample per slice (bilinear)"
        rl_fp = roofline("fp_kernel (two launches: tile classes)", ms["fp_step"], "clinical.fp_step",
                         prof, cal, 8 * samples * S, 8 * samples * S, fpm)
        rl_adj = roofline("fp_adjoint_cell_staged_kernel", ms["fp_adjoint_step"],
                          "clinical.fp_adjoint_step", prof, cal, 8 * samples * S, 8 * samples * S, fpm)
        out["roofline"] = rl_fp if ms["fp_step"] >= ms["fp_adjoint_step"] else rl_adj
        out["roofline_fp"] = rl_fp
        out["roofline_fp_adjoint"] = rl_adj
        bpm = "4 flops and 2 taps of 4 B per voxel-view update (linear interpolation)"
        out["roofline_other"] = {
            k: roofline(kn, ms[k], "clinical." + k, prof, cal, 4 * U * S, 8 * U * S, bpm) for k, kn in (
                ("bp", "bp_staged_kernel"),
                ("fbp", "ramp_gemm_kernel + bp_staged_kernel"),
                ("bp_adjoint", "bp_adjoint_staged_kernel"),
                ("fbp_adjoint", "bp_adjoint_staged_kernel + ramp_gemm_kernel"))}
        result = out
    if args.workload in ("all", "uneven"):
        out, ctx = bench_uneven(torch, args)
        ms, samples = ctx["ms"], ctx["samples"]
        fpm = "8 flops and 4 taps of 2 B per FP sample (bilinear), BF16 storage"
        rl_f = roofline("fp_batch_kernel<8 lanes per ray, bf16, float>", ms["fp_step"], "uneven.fp_step", prof, cal,
                        8 * samples, 8 * samples, fpm)
        rl_a = roofline("fp_adjoint_batch_kernel<bf16, float>", ms["fp_adjoint_step"], "uneven.fp_adjoint_step", prof,
                        cal, 8 * samples, 8 * samples, fpm)
        out["roofline"] = rl_f if ms["fp_step"] >= ms["fp_adjoint_step"] else rl_a
        out["roofline_fp"] = rl_f
        out["roofline_fp_adjoint"] = rl_a
        if result:
            result["uneven"] = out
        else:
            result = dict(out, n_gpus=1, steps=args.steps, warmup=args.warmup, scaling="weak", vs_baseline=None,
                          data="synthetic")
    if args.workload in ("all", "fdk"):
        out, ctx = bench_fdk(torch, args)
        ms, U = ctx["ms"], ctx["U"]
        out["roofline"] = roofline("cone_bp_lds_kernel<32-voxel z-chunks, float>", ms["backproject"],
                                   "fdk.backproject", prof, cal, 8 * U, 16 * U,
                                   "8 flops and 4 taps of 4 B (FP32 filtered projections) per voxel-view update")
        rf = roofline("ramp_gemm_kernel<bf16 rows, cosine pre-weight>", ms["filter"], "fdk.filter", prof, cal,
                      int(2 * ctx["rows"] * ctx["nC"] ** 2), None,
                      "2 rows nCols^2 flops of the dense Toeplitz product (the split operands run three BF16 "
                      "products per output tile, not counted)")
        out["roofline_filter"] = rf
        if result:
            result["fdk"] = out
        else:
            result = dict(out, n_gpus=1, steps=args.steps, warmup=args.warmup, scaling="weak", vs_baseline=None,
                          data="synthetic")
    if cal:
        result["counter_calibration"] = calibration_summary(cal)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
