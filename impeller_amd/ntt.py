"""NTT operators, mirroring the reference's Rust wrappers in core/src/ntt.rs.

Each class exposes get_fn() returning a callable with the reference runtime's
Function signature f(mut_var, var, gpu_mapping, poly_slice)
(runtime/src/functions.rs:20-36), unwrapping the same argument lists the
compiler emits (compiler/src/transit/type3/lowering/emit_func.rs:143-151,
281-308), and calls the C ABI.  Assertions mirror the Rust ones.
"""
import ctypes

import numpy as np

from . import _lib
from .runtime import ScalarArray, Stream


class SsipNtt:
    """core/src/ntt.rs:25-38,161-210: mut_var=[x], var=[twiddle, stream]."""

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 1 and len(var) == 2
            x, twiddle, stream = mut_var[0], var[0], var[1]
            assert x.len & (x.len - 1) == 0, "NTT length must be a power of two"
            log_len = x.len.bit_length() - 1
            assert twiddle.len * 2 == x.len
            _lib.call("ssip_ntt", x.poly_ptr(), twiddle.data_ptr(), log_len, stream.raw(), 8, 8)
        return f


class SsipPrecompute:
    """core/src/ntt.rs:212-245: fills a host twiddle array with unit^i."""

    def get_fn(self):
        def f(twiddle: ScalarArray, unit: np.ndarray):
            assert twiddle.device == "cpu" and twiddle.len & (twiddle.len - 1) == 0
            u = np.ascontiguousarray(unit, dtype=np.uint32)
            _lib.call("ssip_precompute", twiddle.data_ptr(), twiddle.len, u.ctypes.data)
        return f


class RecomputeNtt:
    """core/src/ntt.rs:247-311: mut_var=[x], var=[pq, omegas, stream]."""

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 1 and len(var) == 3
            x, pq, omegas, stream = mut_var[0], var[0], var[1], var[2]
            assert x.len & (x.len - 1) == 0
            log_len = x.len.bit_length() - 1
            assert pq.len & (pq.len - 1) == 0
            pq_deg = (pq.len.bit_length() - 1) + 1
            assert omegas.len == 32
            _lib.call("recompute_ntt", x.poly_ptr(), pq.data_ptr(), pq_deg, omegas.data_ptr(),
                      log_len, stream.raw(), 8, 8)
        return f


class GenPqOmegas:
    """core/src/ntt.rs:313-382 (host tables for RecomputeNtt)."""

    @staticmethod
    def get_deg(deg_stage, max_deg_stage):
        rounds = 1
        while True:
            deg = deg_stage if rounds == 1 else (deg_stage - 1) // rounds + 1
            if deg <= max_deg_stage:
                return deg
            rounds += 1

    def get_pq_len(self, log_len):
        s1, s2 = 8, 8  # get_stage_threads for 32-byte fields (ntt.rs:384-389)
        d1 = self.get_deg((log_len + 1) // 2, s1 + 1)
        d2 = self.get_deg(log_len // 2, (s2 + 2) // 2)
        return 1 << (max(d1, d2) - 1)

    def get_fn(self):
        def f(pq: ScalarArray, omegas: ScalarArray, length: int, unit: np.ndarray):
            assert omegas.len == 32
            pq_deg = (pq.len.bit_length() - 1) + 1
            u = np.ascontiguousarray(unit, dtype=np.uint32)
            _lib.lib().gen_pq_omegas(pq.data_ptr(), omegas.data_ptr(), pq_deg, length,
                                     u.ctypes.data)
        return f


class DistributePowers:
    """core/src/ntt.rs:108-159: mut_var=[x], var=[powers, stream], slice-aware."""

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 1 and len(var) == 2
            x, powers, stream = mut_var[0], var[0], var[1]
            if poly_slice is None:
                poly_slice = _lib.PolySlice(0, x.len, x.len)
            _lib.call("distribute_powers", x.poly_ptr(), powers.data_ptr(), powers.len,
                      stream.raw(), ctypes.byref(poly_slice))
        return f


# ------------------------------------------------------ device-resident API
def twiddles(omega: np.ndarray, n: int, stream: Stream, device_id: int = 0):
    """Device table omega^i, i < n (a torch int32 tensor of n*8 words)."""
    from .runtime import device_u32
    t = device_u32(max(n, 1) * 8, device_id)
    o = np.ascontiguousarray(omega, dtype=np.uint32)
    _lib.call("impeller_twiddles", t.data_ptr(), n, o.ctypes.data, stream.raw())
    return t


def ntt(x_tensor, twiddle_tensor, log_len, stream: Stream, inverse=False, coset_shift=None,
        passes=None):
    """In-place NTT of a contiguous device tensor (impeller_ntt).

    passes: optional plan [fuse, lds_log, d0, d1, ...] -- levels fused into the
    bit-reversal kernel (0, 4 or 5; or 1: no bit-reversal kernel, the first
    pass reads the input bit-reversed into a second buffer, at least two
    passes), log2 elements staged per work-group (10 or 11) and the level
    count of each following pass (levels of fuse + sum(d) == log_len).
    """
    plan = None
    n_passes = 0
    if passes:
        arr = (ctypes.c_uint * len(passes))(*passes)
        plan = ctypes.cast(arr, ctypes.c_void_p)
        n_passes = len(passes)
    shift = coset_shift.data_ptr() if coset_shift is not None else None
    _lib.call("impeller_ntt", x_tensor.data_ptr(), twiddle_tensor.data_ptr(), log_len,
              1 if inverse else 0, shift, plan, n_passes, stream.raw())


# ------------------------------------------------------------- batched NTT
class NttBatch:
    """A prepared batch of independent NTTs (impeller_ntt_batch_*).

    items: dicts with keys data (device tensor, 2^log_n elements, result in
    place), log_n, twiddle (device table of the omega used), and optionally
    src (device tensor, 2^log_src coefficients, zero-extended), log_src,
    inverse, coset_shift (device element), plan (list).
    """

    def __init__(self, items, stream: Stream):
        from .runtime import device_u32
        n = len(items)
        arr = (_lib.ImpNttItem * n)()
        self._keep = []
        for i, it in enumerate(items):
            a = arr[i]
            a.data = it["data"].data_ptr()
            a.src = it["src"].data_ptr() if it.get("src") is not None else None
            a.log_n = it["log_n"]
            a.log_src = it.get("log_src", it["log_n"])
            a.twiddle = it["twiddle"].data_ptr()
            a.inverse = 1 if it.get("inverse") else 0
            a.coset_shift = it["coset_shift"].data_ptr() if it.get("coset_shift") is not None else None
            plan = it.get("plan") or []
            a.plan_len = len(plan)
            for k, v in enumerate(plan):
                a.plan[k] = v
            self._keep.append(it)
        # room for the second buffer of every fused-reversal item
        nbytes = int(_lib.call("impeller_ntt_batch_bytes_items", arr, n))
        self.dev = device_u32((nbytes + 3) // 4)
        self.info = _lib.ImpNttBatch()
        _lib.call("impeller_ntt_batch_prepare", arr, n, self.dev.data_ptr(), nbytes,
                  ctypes.byref(self.info), stream.raw())

    def run(self, stream: Stream, max_blocks=0):
        _lib.call("impeller_ntt_batch_run", ctypes.byref(self.info), max_blocks, stream.raw())


class Graph:
    """A HIP graph captured from the work a callable launches on stream."""

    def __init__(self, stream: Stream, fn):
        _lib.call("impeller_graph_begin", stream.raw())
        try:
            fn()
        finally:
            h = ctypes.c_void_p()
            _lib.call("impeller_graph_end", stream.raw(), ctypes.byref(h))
        self.handle = h

    def launch(self, stream: Stream):
        _lib.call("impeller_graph_launch", self.handle, stream.raw())

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().impeller_graph_destroy(self.handle)
        except Exception:
            pass
