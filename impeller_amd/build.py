"""Builds libimpeller_hip.so (gfx950) in-tree with hipcc.

Usage: python -m impeller_amd.build   (or impeller_amd.build.build())
Each .hip source under csrc/ is compiled to an object in parallel, then linked
into impeller_amd/libimpeller_hip.so next to this file.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libimpeller_hip.so")
OBJDIR = os.path.join(HERE, "_build")
ARCH = os.environ.get("IMPELLER_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep hipcc from emitting packed-fp32 instructions (none of this code
    # is floating point except the cost model, which stays scalar)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-munsafe-fp-atomics",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


READELF = os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "llvm", "bin",
                       "llvm-readelf")
if not os.path.exists(READELF):
    READELF = shutil.which("llvm-readelf") or "/opt/rocm/llvm/bin/llvm-readelf"


def _device_functions(src, include_dirs=()):
    """Out-of-line device functions (FUNC symbols without a kernel descriptor)
    in the gfx950 code object of one source.  There must be none: in a callee
    the compiler's long-branch expansion has been seen to reuse the return
    address registers s[30:31] without saving them, so a wave that takes such
    a branch never returns.  Every device function here is force-inlined and
    this check keeps it that way."""
    obj = os.path.join(OBJDIR if not include_dirs else os.path.dirname(src),
                       os.path.basename(src) + ".dev.o")
    inc = [a for d in include_dirs for a in ("-I", d)]
    cmd = [HIPCC, *COMMON, *inc, "--cuda-device-only", "--no-gpu-bundle-output", "-c", src,
           "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc (device object) failed for {src}:\n{r.stderr}")
    r = subprocess.run([READELF, "-s", "--wide", obj], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"llvm-readelf failed on {obj}:\n{r.stderr}")
    funcs, kds = set(), set()
    for line in r.stdout.splitlines():
        f = line.split()
        if len(f) >= 8 and f[3] == "FUNC":
            funcs.add(f[7])
        elif len(f) >= 8 and f[3] == "OBJECT" and f[7].endswith(".kd"):
            kds.add(f[7][:-3])
    return sorted(funcs - kds)


def build(verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    workers = min(2 * len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        objs_f = [ex.submit(_compile, s) for s in srcs]
        calls_f = [ex.submit(_device_functions, s) for s in srcs]
        objs = [f.result() for f in objs_f]
        calls = {s: f.result() for s, f in zip(srcs, calls_f)}
    bad = {os.path.basename(s): c for s, c in calls.items() if c}
    if bad:
        raise RuntimeError(f"out-of-line device functions (force-inline them): {bad}")
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
