"""Builds the library's HIP sources for the CPU emulation (hip/hip_runtime.h,
emu.cpp) into tests/emu/_build/libimpeller_emu[_asan].so.

  python tests/emu/build.py [--asan]

The sources are compiled as plain C++ for the host (the default, compiler-
scheduled Montgomery multiply); every kernel runs through the emulated launch, so the
same C ABI can be driven from Python with host buffers standing in for device
memory.  Test infrastructure only: the product library is impeller_amd/build.py.
"""
import concurrent.futures
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "impeller_amd", "csrc")
OUT_DIR = os.path.join(HERE, "_build")
CLANG = "/opt/rocm/llvm/bin/clang++" if os.path.exists("/opt/rocm/llvm/bin/clang++") \
    else (shutil.which("clang++") or "clang++")


def lib_path(asan=False):
    return os.path.join(OUT_DIR, "libimpeller_emu_asan.so" if asan else "libimpeller_emu.so")


# the runtime layer calls the HIP runtime itself (virtual memory, pinned
# memory, events): it has no kernels and is not emulated
NOT_EMULATED = ("runtime_api.hip",)


def sources():
    return sorted(p for p in glob.glob(os.path.join(CSRC, "*.hip"))
                  if os.path.basename(p) not in NOT_EMULATED) + [os.path.join(HERE, "emu.cpp")]


def build(asan=False, verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    flags = ["-std=c++17", "-O1", "-g", "-fPIC", "-I", HERE, "-DIMP_EMU",
             "-Wno-unused-result", "-Wno-unknown-attributes"]
    if asan:
        flags += ["-fsanitize=address", "-fno-omit-frame-pointer"]
    tag = "asan" if asan else "plain"

    def one(src):
        obj = os.path.join(OUT_DIR, f"{os.path.basename(src)}.{tag}.o")
        if os.path.exists(obj) and os.path.getmtime(obj) > max(
                os.path.getmtime(p) for p in glob.glob(os.path.join(CSRC, "*"))
                + glob.glob(os.path.join(HERE, "hip", "*")) + [src]):
            return obj
        cmd = [CLANG, *flags, "-x", "c++", "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"emu compile failed for {src}:\n{r.stderr[-4000:]}")
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=6) as ex:
        objs = list(ex.map(one, sources()))
    out = lib_path(asan)
    cmd = [CLANG, "-shared", "-fPIC", *(["-fsanitize=address"] if asan else []), *objs,
           "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emu link failed:\n{r.stderr[-4000:]}")
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {out}")
    return out


def build_extra(srcs, out, include_dirs=()):
    """Compiles extra HIP sources (e.g. emitted fused kernels) with the
    emulation runtime into their own shared object (its own copy of the
    launch emulation; kernels launched from it run on the host)."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    flags = ["-std=c++17", "-O1", "-fPIC", "-I", HERE, "-DIMP_EMU", "-Wno-unused-result",
             "-Wno-unknown-attributes", "-I", CSRC]
    for d in include_dirs:
        flags += ["-I", d]
    objs = []
    for src in list(srcs) + [os.path.join(HERE, "emu.cpp")]:
        obj = os.path.join(os.path.dirname(out), os.path.basename(src) + ".emu.o")
        r = subprocess.run([CLANG, *flags, "-x", "c++", "-c", src, "-o", obj],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"emu compile failed for {src}:\n{r.stderr[-4000:]}")
        objs.append(obj)
    r = subprocess.run([CLANG, "-shared", "-fPIC", *objs, "-o", out], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emu link failed:\n{r.stderr[-4000:]}")
    return out


def asan_runtime():
    r = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"],
                       capture_output=True, text=True)
    return r.stdout.strip()


if __name__ == "__main__":
    build(asan="--asan" in sys.argv, verbose=True)
