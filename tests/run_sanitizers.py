"""Host AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU tests.

    python tests/run_sanitizers.py

Builds libfbct.so with its host code instrumented (hipcc, the sanitizers on the host
compilation only: GPU code is never built with them) and the FP64 oracle with clang, both
into ``_build_san/`` directories, then runs the CPU tests that exercise them --
tests/test_host.py (plan creation, FP64 tables, argument validation, uneven-batch layout,
overlap checks) and the oracle tests -- in a child Python with the clang ASan runtime
preloaded (Python itself is not instrumented, so leak checks are off).  Any sanitizer
report aborts the child and fails the run.  CPU only: nothing here touches a GPU.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "fanbeamctrecon_tensorflowop_amd")
OUT = os.path.join(PKG, "_build_san")
ORACLE_OUT = os.path.join(REPO, "oracle", "_build_san")
ROCM = "/opt/rocm"
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
CLANG = os.path.join(ROCM, "llvm", "bin", "clang")
# every sanitizer flag applies to the host compilation only
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]


def asan_runtime() -> str:
    libs = sorted(glob.glob(os.path.join(ROCM, "llvm", "lib", "clang", "*", "lib", "linux",
                                         "libclang_rt.asan-x86_64.so")))
    if not libs:
        raise SystemExit("clang ASan runtime not found under %s/llvm" % ROCM)
    return libs[-1]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise SystemExit("failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout[-2000:], r.stderr[-4000:]))


def build_library() -> str:
    sys.path.insert(0, REPO)
    from fanbeamctrecon_tensorflowop_amd import build as b
    os.makedirs(OUT, exist_ok=True)

    def one(src):
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        flags = b.COMMON + (b.DEVICE if src.endswith(".hip") else b.HOST) + SAN
        run([HIPCC] + flags + ["-c", src, "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(one, b.sources()))
    lib = os.path.join(OUT, "libfbct.so")
    run([HIPCC, "--offload-arch=" + b.ARCH, "-shared", "-fPIC", "-o", lib] + objs + SAN +
        ["-Xarch_host", "-shared-libasan", "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64"])
    return lib


def build_oracle() -> str:
    os.makedirs(ORACLE_OUT, exist_ok=True)
    lib = os.path.join(ORACLE_OUT, "libfbct_oracle.so")
    run([CLANG, "-O1", "-g", "-std=c11", "-fPIC", "-pthread", "-ffp-contract=off", "-shared",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-shared-libasan",
         "-o", lib, os.path.join(REPO, "oracle", "fbct_oracle.c"), "-lm"])
    return lib


def main() -> int:
    lib, orc = build_library(), build_oracle()
    env = dict(os.environ)
    pre = [asan_runtime()] + ([env["LD_PRELOAD"]] if env.get("LD_PRELOAD") else [])
    env.update({
        "LD_PRELOAD": ":".join(pre),
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=1:detect_odr_violation=0",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "FBCT_LIBRARY": lib,
        "FBCT_ORACLE_LIBRARY": orc,
    })
    tests = ["tests/test_host.py", "tests/test_oracle.py", "tests/test_oracle_cone.py", "tests/test_golden.py"]
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"] + tests
    print("sanitized run:", " ".join(tests), flush=True)
    r = subprocess.run(cmd, cwd=REPO, env=env)
    print("host ASan+UBSan run: %s" % ("passed" if r.returncode == 0 else "FAILED (rc %d)" % r.returncode),
          flush=True)
    return r.returncode


if __name__ == "__main__":
    sys.exit(main())
