"""Build libfbct.so (HIP kernels + C ABI host library) for gfx950, in-tree.

    python -m fanbeamctrecon_tensorflowop_amd.build [--jobs N]

Every source under csrc/ is compiled with hipcc for gfx950 and linked into
``fanbeamctrecon_tensorflowop_amd/libfbct.so``.  No CMake, no JIT cache.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
BUILD_DIR = os.path.join(PKG_DIR, "_build")
LIB_PATH = os.path.join(PKG_DIR, "libfbct.so")
PROBE_SRC = os.path.join(PKG_DIR, "probes", "probes.hip")
PROBE_LIB_PATH = os.path.join(PKG_DIR, "libfbct_probes.so")
ARCH = os.environ.get("FBCT_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(REPO, "include")]
# Device code: no packed FP32 instructions (v_pk_fma_f32 & co.) are generated.
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
# Host-only geometry: no contraction, so the FP64 tables follow the written expression order.
HOST = ["-x", "c++", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    flags = COMMON + (DEVICE if src.endswith(".hip") else HOST)
    cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    # hipcc's host pass does not know the device target feature; drop that one warning
    noise = [ln for ln in (r.stderr or "").splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "warning generated" not in ln]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")
    return obj


def _link(objs, path: str) -> None:
    tmp = path + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + [
        "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, path)


def build(jobs: int = 0, verbose: bool = False) -> str:
    """Build libfbct.so and the counter-calibration probes (libfbct_probes.so, measurement
    only: bench.py's profile mode loads it, the operators never do)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sources() + [PROBE_SRC]
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    _link(objs[:-1], LIB_PATH)
    _link(objs[-1:], PROBE_LIB_PATH)
    if verbose:
        print("built", LIB_PATH, "and", PROBE_LIB_PATH)
    return LIB_PATH


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    a = ap.parse_args(argv)
    build(a.jobs, verbose=True)


if __name__ == "__main__":
    main()
