"""Loads the CPU-emulated library (tests/emu/build.py) with the product's
ctypes signatures; numpy arrays stand in for device buffers.  Used by
tests/test_emu_kernels.py and runnable under AddressSanitizer:

  LD_PRELOAD=$(clang++ -print-file-name=libclang_rt.asan-x86_64.so) \
  ASAN_OPTIONS=detect_leaks=0 python tests/emu/harness.py --asan [scenario ...]
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from impeller_amd import _lib  # noqa: E402  (signatures and struct mirrors only)

_L = {}


def load(asan=False):
    if asan not in _L:
        from tests.emu import build as B
        path = B.lib_path(asan)
        if not os.path.exists(path):
            B.build(asan=asan)
        L = ctypes.CDLL(path)
        for name, (res, args) in _lib.SIGNATURES.items():
            if name.startswith("impeller_rt_"):
                continue  # the runtime layer is not emulated (tests/emu/build.py)
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _L[asan] = L
    return _L[asan]


class Emu:
    def __init__(self, asan=False):
        self.L = load(asan)

    def call(self, name, *args):
        r = getattr(self.L, name)(*args)
        if _lib.SIGNATURES[name][0] is _lib._ERR and r != 0:
            raise _lib.HipError(name, r)
        return r


def ptr(a):
    return None if a is None else a.ctypes.data


def u32(n):
    return np.zeros(n, np.uint32)


def garbage_u32(n, fill=0xA5A5A5A5):
    """A buffer with non-zero contents, like freshly allocated device memory."""
    return np.full(n, fill, np.uint32)
