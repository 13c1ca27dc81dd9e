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
