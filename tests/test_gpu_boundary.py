"""The reference-ABI contract on the GPU: ssip_ntt / recompute_ntt allocate
nothing and match best_fft (in-place plans 2^20..2^24, recompute twiddles
from pq / omegas), the generated-kernel gate reads its views from device
memory and is captured into a HIP graph, and the host-buffer msm streams
the points through part buffers (closed form)."""
import ctypes
import random

import numpy as np
import pytest
import torch

from impeller_amd import _lib
from impeller_amd import ntt as N
from impeller_amd import workloads as WL
from impeller_amd.fused import FusedGate
from impeller_amd.runtime import ScalarArray, device_u32, to_device, to_host
from oracle import oracle as O

pytestmark = pytest.mark.gpu
mont = WL.mont_limbs
R = WL.R_MOD


def _free_bytes():
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


@pytest.mark.parametrize("k", [20, 21, 22, 23])
def test_ssip_ntt_inplace_plans(stream, k):
    """ssip_ntt's in-place plan of each size (rotated view) against best_fft."""
    n = 1 << k
    x = O.gen_scalars(n, 3000 + k)
    w = mont(random.Random(k).randrange(R))  # formal omega (core/tests/ntt.rs:63)
    expect = O.best_fft(x, w, k)
    rot = 12345 % n
    xa = ScalarArray.from_numpy(np.roll(x, -rot, axis=0))
    xa.rotate(rot)
    tw = ScalarArray.from_numpy(O.powers(w, n // 2))
    N.SsipNtt().get_fn()([xa], [tw, stream])
    stream.sync()
    assert (xa.logical_numpy() == expect).all()


@pytest.mark.parametrize("k,random_omega", [(10, True), (16, True), (20, False), (22, True)])
def test_recompute_ntt(stream, k, random_omega):
    """recompute_ntt with the pq length the reference wrapper computes
    (GenPqOmegas.get_pq_len) against best_fft; tables from gen_pq_omegas."""
    n = 1 << k
    x = O.gen_scalars(n, 3100 + k)
    w = mont(random.Random(100 + k).randrange(R)) if random_omega else mont(WL.omega(k))
    expect = O.best_fft(x, w, k)
    g = N.GenPqOmegas()
    pq = ScalarArray(g.get_pq_len(k), device="cpu")
    om = ScalarArray(32, device="cpu")
    g.get_fn()(pq, om, n, w)
    pq_d = ScalarArray.from_numpy(pq.to_numpy())
    om_d = ScalarArray.from_numpy(om.to_numpy())
    xa = ScalarArray.from_numpy(x)
    N.RecomputeNtt().get_fn()([xa], [pq_d, om_d, stream])
    stream.sync()
    assert (xa.to_numpy() == expect).all()


def test_reference_ntts_allocate_nothing(stream):
    """ssip_ntt and recompute_ntt at 2^22 leave the device's free memory
    unchanged (hipMemGetInfo before and after)."""
    k = 22
    n = 1 << k
    x = device_u32(n * 8)
    _lib.call("impeller_gen_scalars", x.data_ptr(), n, 0, 5, 0, stream.raw())
    w = mont(WL.omega(k))
    tw = N.twiddles(w, n // 2, stream)
    g = N.GenPqOmegas()
    pq = ScalarArray(g.get_pq_len(k), device="cpu")
    om = ScalarArray(32, device="cpu")
    g.get_fn()(pq, om, n, w)
    pq_d, om_d = to_device(pq.to_numpy()), to_device(om.to_numpy())
    pp = _lib.PolyPtr(x.data_ptr(), n, 0, 0, n)
    before = _free_bytes()
    for _ in range(3):
        _lib.call("ssip_ntt", pp, tw.data_ptr(), k, stream.raw(), 8, 8)
        _lib.call("recompute_ntt", pp, pq_d.data_ptr(), (pq.len.bit_length() - 1) + 1,
                  om_d.data_ptr(), k, stream.raw(), 8, 8)
    after = _free_bytes()
    assert after == before, (before, after)


def test_fused_gate_graph_capture(stream):
    """fused_gate_abc_d_qe (views in a device array, rotated input and output)
    captured into a HIP graph and replayed with eager work between replays."""
    n = 1 << 16
    ins = [O.gen_scalars(n, 70 + i) for i in range(6)]
    expect = O.gate(*ins)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        from impeller_amd.runtime import Stream
        st = Stream.current(0)
        arrs = []
        for i, a in enumerate(ins):
            rot = 7 * i
            sa = ScalarArray.from_numpy(np.roll(a, -rot, axis=0))
            if rot:
                sa.rotate(rot)
            arrs.append(sa)
        out = ScalarArray(n, device="cuda")
        out.rotate(3)
        arg = device_u32(7 * 10)
        fn = FusedGate().get_fn()
        fn([arg, None, out], [st] + arrs)  # eager, also fills the arg buffer
        st.sync()
        assert (out.logical_numpy() == expect).all()
        # the graph holds the kernel launch only (the arg buffer was filled
        # by the eager call, as the runtime fills it before the launch)
        base = arg.data_ptr()
        sl = _lib.PolySlice(0, n, n)
        g = N.Graph(st, lambda: _lib.call("fused_gate_abc_d_qe", base, base + 6 * 40, n,
                                          ctypes.byref(sl), None, st.raw()))
        for rep in range(3):
            out.values.fill_(-1)  # eager work between replays
            torch.cuda.synchronize()
            g.launch(st)
            st.sync()
            assert (out.logical_numpy() == expect).all(), rep


@pytest.mark.parametrize("parts,stage_points,stage_scalers,bpr,batches",
                         [(4, 2, 2, 1, 2), (1, 1, 1, 2, 3)])
def test_msm_host_pipeline_22(stream, parts, stage_points, stage_scalers, bpr, batches):
    """The reference's host-buffer msm at 2^22 (window 16, no tables): pinned
    host points and scalars, `parts` ranges through `stage_points` buffers;
    each batch's affine result against the closed form.  With parts = 4 the
    device buffer shrinks with the part size."""
    nm = 1 << 22
    bases = device_u32(nm * 16)
    _lib.call("impeller_gen_bases", bases.data_ptr(), nm, 77, 0, stream.raw())
    pts_h = torch.empty(nm * 16, dtype=torch.int32, pin_memory=True)
    pts_h.copy_(bases)
    del bases
    kk = O.gen_scalars(nm, 77)
    scal_h, expects = [], []
    for b in range(batches):
        s = O.gen_scalars(nm, 4400 + b, kind=b % 3)
        t = torch.empty(nm * 8, dtype=torch.int32, pin_memory=True)
        t.copy_(torch.from_numpy(s.reshape(-1).view(np.int32)))
        scal_h.append(t)
        expects.append(O.msm_closed_form(s, kk))
    res_h = [torch.zeros(16, dtype=torch.int32, pin_memory=True) for _ in range(batches)]
    sizes = (ctypes.c_ulong * 1)()
    cards = (ctypes.c_uint * 1)(0)
    _lib.call("msm", None, sizes, nm, bpr, parts, stage_scalers, stage_points, 1, cards, None,
              0, None, None)
    if parts > 1:
        one = (ctypes.c_ulong * 1)()
        _lib.call("msm", None, one, nm, bpr, 1, stage_scalers, 1, 1, cards, None, 0, None, None)
        assert sizes[0] < one[0], "parts should shrink the device buffer"
    buf = device_u32((sizes[0] + 3) // 4)
    arr = lambda xs: (ctypes.c_void_p * len(xs))(*[x.data_ptr() for x in xs])  # noqa: E731
    _lib.call("msm", arr([buf]), None, nm, bpr, parts, stage_scalers, stage_points, 1, cards,
              arr([pts_h]), batches, arr(scal_h), arr(res_h))
    for b in range(batches):
        got = res_h[b].numpy().view(np.uint32)
        assert (got == expects[b]).all(), b
