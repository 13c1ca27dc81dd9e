"""C-ABI boundary checks that need no GPU: the library loads, exports every
symbol include/impeller_hip.h declares, and the host-side mirror of the
reference's operator types behaves like the reference."""
import ctypes
import os
import re

import numpy as np

from impeller_amd import _lib
from impeller_amd.msm import MsmConfig
from impeller_amd.ntt import GenPqOmegas
from impeller_amd.runtime import ScalarArray

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "impeller_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


def test_header_symbols_exported():
    syms = declared_symbols()
    assert "ssip_ntt" in syms and "msm" in syms and "fused_gate_abc_d_qe" in syms
    L = ctypes.CDLL(_lib.LIB_PATH)
    for s in syms:
        assert hasattr(L, s), f"{s} declared in impeller_hip.h but not exported"
    # the python binding declares exactly the header's symbols
    assert set(_lib.SIGNATURES) == set(syms)


def test_struct_layouts():
    # core/src/poly_ptr.rs:4-11: pointer + 4 usize = 40 bytes; PolySlice 24 bytes
    assert ctypes.sizeof(_lib.PolyPtr) == 40
    assert ctypes.sizeof(_lib.PolySlice) == 24


def test_version_and_bad_args_without_gpu():
    assert b"gfx950" in _lib.lib().impeller_version()
    # argument validation happens before any device work
    L = _lib.lib()
    p = _lib.PolyPtr(None, 8, 0, 0, 8)
    assert L.ssip_ntt(p, None, 3, None, 8, 8) != 0          # null pointers
    p = _lib.PolyPtr(1234, 6, 0, 0, 6)
    assert L.ssip_ntt(p, 1234, 3, None, 8, 8) != 0          # len != 2^log_len
    assert L.impeller_msm_workspace_size(1 << 10, 2) == 0   # window out of range
    assert L.impeller_msm_workspace_size(1 << 10, 16) > 0
    sizes = (ctypes.c_ulong * 2)()
    cards = (ctypes.c_uint * 2)(0, 1)
    # buffer-size query protocol (core/src/msm.rs:282-302)
    assert L.msm(None, sizes, 1 << 16, 1, 1, 1, 1, 2, cards, None, 0, None, None) == 0
    assert sizes[0] > (1 << 16) * 96 and sizes[1] == 0


def test_scalar_array_views_match_runtime_semantics():
    # runtime/src/scalar.rs:438-477: slice of a rotated array records the
    # physical start (len + start - rotate) % len
    a = ScalarArray(16, np.arange(16 * 8, dtype=np.uint32).reshape(16, 8), "cpu")
    a.rotate(3)
    s = a.slice(4, 12)
    assert s.slice_info.offset == (16 + 4 - 3) % 16 and s.slice_info.whole_len == 16
    p = s.poly_ptr()
    assert (p.len, p.rotate, p.offset, p.whole_len) == (8, 0, 1, 16)
    # logical element 0 of the rotated array is physical len - rotate
    assert (a.logical_numpy()[0] == a.to_numpy()[13]).all()
    b = ScalarArray(16, None, "cpu")
    b.rotate(-2)
    assert b.get_rotation() == 14


def test_msm_config_matches_reference_formula():
    # common/src/msm_config.rs:17-26 with the library's window count
    for c in range(8, 21):
        for t in (1, 2, 4, 16, 64):
            cfg = MsmConfig(window_size=c, target_window=t)
            a = 254 // c + 1
            nw = min(a, t)
            assert cfg.get_precompute() == -(-a // nw)


def test_pq_len_matches_reference():
    # core/src/ntt.rs:333-346 with (8, 8) stage threads
    g = GenPqOmegas()
    # k=20: stage degrees get_deg(10, 9) = 5 and get_deg(10, 5) = 5 -> 2^4
    assert g.get_pq_len(20) == 16
    # k=24: get_deg(12, 9) = 6, get_deg(12, 5) = 4 -> 2^5
    assert g.get_pq_len(24) == 32


def _fe(x):
    from oracle import oracle as O
    return O.int_to_limbs(O.to_mont_int(x))


def _pow_fe(base_int, e):
    from oracle import oracle as O
    return _fe(pow(base_int, e, O.R_MOD))


def test_gen_pq_omegas_host_values():
    """gen_pq_omegas runs on the host (precompute.cuh:89-112): omegas[i] =
    unit^(2^i) always; pq[0] = 1, pq[i] = unit^(i * (len >> pq_deg)), also
    for pq_deg 0 / 1 and a length that is not a power of two."""
    import random
    from oracle import oracle as O
    L = _lib.lib()
    rng = random.Random(5)
    for pq_deg, length in ((0, 1 << 10), (1, 1 << 10), (2, 12), (5, 3 << 10), (9, 1 << 20)):
        u = rng.randrange(1, O.R_MOD)
        npq = max(1, (1 << pq_deg) >> 1)
        pq = np.full((npq, 8), 0xDEADBEEF, np.uint32)
        om = np.full((32, 8), 0xDEADBEEF, np.uint32)
        unit = _fe(u)
        L.gen_pq_omegas(pq.ctypes.data, om.ctypes.data, pq_deg, length, unit.ctypes.data)
        for i in range(32):
            assert (om[i] == _pow_fe(u, 1 << i)).all(), (pq_deg, i)
        step = length >> pq_deg
        for i in range(npq):
            assert (pq[i] == _pow_fe(u, i * step)).all(), (pq_deg, length, i)
    # null pointers leave the outputs alone
    om = np.full((32, 8), 7, np.uint32)
    L.gen_pq_omegas(None, om.ctypes.data, 3, 64, _fe(3).ctypes.data)
    assert (om == 7).all()


def test_msm_rejects_point_index_overflow():
    """Point indices are packed into 31 bits: n_pre * len >= 2^31 is an
    argument error before any device work (host ABI and device ABI)."""
    L = _lib.lib()
    sizes = (ctypes.c_ulong * 1)()
    cards = (ctypes.c_uint * 1)(0)
    assert L.msm(None, sizes, 1 << 31, 1, 1, 1, 1, 1, cards, None, 0, None, None) != 0
    # parts split the points: each part's indices fit again
    assert L.msm(None, sizes, 1 << 31, 1, 4, 1, 1, 1, cards, None, 0, None, None) == 0
    fake = ctypes.c_void_p(4096)
    assert L.impeller_msm(fake, fake, 1 << 31, 16, fake, fake, 1 << 40, None) != 0


def test_msm_host_buffer_shrinks_with_parts():
    L = _lib.lib()
    sizes = (ctypes.c_ulong * 1)()
    cards = (ctypes.c_uint * 1)(0)
    got = []
    for parts, sp in ((1, 1), (4, 2), (16, 2), (64, 2)):
        assert L.msm(None, sizes, 1 << 22, 1, parts, 1, sp, 1, cards, None, 0, None, None) == 0
        got.append(sizes[0])
    # the staged points and scalars and the sort workspace (16 bytes per
    # point and window) shrink with the part; the bucket arrays and the
    # bucket sums kept across parts (per window, independent of len) stay
    assert got[0] > (1 << 22) * 64 and got[1] < got[0] and got[2] < got[1] and got[3] < got[2]
    assert got[3] < got[0] // 6


def test_ntt_workspace_size():
    L = _lib.lib()
    fused = (ctypes.c_uint * 5)(1, 10, 8, 8, 8)
    inplace = (ctypes.c_uint * 5)(0, 10, 10, 8, 6)
    base = L.impeller_ntt_workspace_size(24, inplace, 5)
    assert base > 0
    assert L.impeller_ntt_workspace_size(24, fused, 5) == base + (32 << 24)
    bad = (ctypes.c_uint * 4)(1, 10, 8, 8)  # levels do not add up to 24
    assert L.impeller_ntt_workspace_size(24, bad, 4) == 0


def test_runtime_layer_rejects_bad_arguments_without_gpu():
    """impeller_rt_* check their arguments before any runtime call and
    return an error where the reference's crate would panic."""
    L = _lib.lib()
    out = ctypes.c_void_p()
    assert L.impeller_rt_static_allocate(None, 0, 16, ctypes.byref(out)) != 0
    assert L.impeller_rt_memcpy_async(0, None, None, 16, 7, None) != 0   # bad copy kind
    ids = (ctypes.c_size_t * 1)(0)
    assert L.impeller_rt_page_allocate(None, 1 << 21, ids, 1, 0, ctypes.byref(out)) != 0
    assert L.impeller_rt_event_sync(None) != 0
    assert L.impeller_rt_event_raw(None) is None
    assert L.impeller_rt_page_size(None) == 0
    assert L.impeller_rt_file_read_to_device(-1, 0, 16, None, None, 0, None) != 0
    # the Python mirror of the crate imports without a GPU
    from impeller_amd import hip_api
    assert hip_api.H2D == 1 and hip_api.D2H == 2 and hip_api.D2D == 3

