"""Fused gate kernel vs the oracle, through the reference's fused-kernel
calling convention (core/src/fused_kernels.rs:95-200)."""
import numpy as np
import pytest

from impeller_amd.fused import FusedGate
from impeller_amd.runtime import ScalarArray, device_u32
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 1000, 1 << 16])
def test_gate(stream, n):
    ins = [O.gen_scalars(n, 10 + i) for i in range(6)]
    expect = O.gate(*ins)
    xs = [ScalarArray.from_numpy(a) for a in ins]
    out = ScalarArray.from_numpy(np.zeros((n, 8), np.uint32))
    arg = device_u32(7 * 10)
    FusedGate().get_fn()([arg, None, out], [stream] + xs)
    stream.sync()
    assert (out.to_numpy() == expect).all()


def test_gate_rotated_views(stream):
    n = 4096
    ins = [O.gen_scalars(n, 20 + i) for i in range(6)]
    expect = O.gate(*ins)
    xs = []
    for i, a in enumerate(ins):
        arr = ScalarArray.from_numpy(np.zeros((n, 8), np.uint32))
        arr.rotate(i * 5 + 1)
        phys = (n - arr.get_rotation() + np.arange(n)) % n
        whole = np.zeros((n, 8), np.uint32)
        whole[phys] = a
        arr.values.copy_(ScalarArray.from_numpy(whole).values)
        xs.append(arr)
    out = ScalarArray.from_numpy(np.zeros((n, 8), np.uint32))
    arg = device_u32(7 * 10)
    FusedGate().get_fn()([arg, None, out], [stream] + xs)
    stream.sync()
    assert (out.to_numpy() == expect).all()
