"""impeller_ntt_noalloc (scratch in the library's static device array)
against impeller_ntt through the C ABI: bit-identical results on two calls in
a row, the second smaller, for coset-forward and inverse-coset transforms,
with no change in the device's free memory."""
import ctypes

import numpy as np
import pytest
import torch

from impeller_amd import _lib
from impeller_amd import ntt as N
from impeller_amd import workloads as WL
from impeller_amd.runtime import device_u32, to_device

pytestmark = pytest.mark.gpu
mont = WL.mont_limbs
R = WL.R_MOD


def _plan(p):
    if not p:
        return None, 0
    arr = (ctypes.c_uint * len(p))(*p)
    return ctypes.cast(arr, ctypes.c_void_p), len(p)


def test_noalloc_matches_impeller_ntt(stream):
    cases = [
        # (log size, inverse, coset shift, plan): the second call is smaller
        (22, 0, WL.GENERATOR, [1, 10, 8, 7, 7]),
        (16, 1, pow(WL.GENERATOR, -1, R), None),
    ]
    for k, inverse, shift, plan in cases:
        n = 1 << k
        x = device_u32(n * 8)
        _lib.call("impeller_gen_scalars", x.data_ptr(), n, 0, 800 + k, 0, stream.raw())
        tw = N.twiddles(mont(WL.omega(k, bool(inverse))), n // 2, stream)
        sh = to_device(mont(shift))
        a, b = x.clone(), x.clone()
        pl, npl = _plan(plan)
        _lib.call("impeller_ntt", a.data_ptr(), tw.data_ptr(), k, inverse, sh.data_ptr(), pl,
                  npl, stream.raw())
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        _lib.call("impeller_ntt_noalloc", b.data_ptr(), tw.data_ptr(), k, inverse,
                  sh.data_ptr(), pl, npl, stream.raw())
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] >= free0, k  # nothing allocated
        assert torch.equal(a, b), (k, inverse)
        assert not torch.equal(b, x), k  # the transform did run


def test_noalloc_rejects_oversize():
    L = _lib.lib()
    fake = ctypes.c_void_p(4096)
    assert L.impeller_ntt_noalloc(fake, fake, 32, 0, None, None, 0, None) != 0
