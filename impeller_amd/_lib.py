"""ctypes binding of libimpeller_hip.so (the C ABI in include/impeller_hip.h).

The library is required: importing a function from here on a machine where
the shared object is missing raises, there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libimpeller_hip.so")


class HipError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} returned hipError {code}")
        self.code = code


class PolyPtr(ctypes.Structure):
    """core/src/poly_ptr.rs:4-11 / core/src/common/iter/src/iter.cuh:12-18"""
    _fields_ = [("ptr", ctypes.c_void_p), ("len", ctypes.c_size_t),
                ("rotate", ctypes.c_size_t), ("offset", ctypes.c_size_t),
                ("whole_len", ctypes.c_size_t)]


ConstPolyPtr = PolyPtr  # identical layout (iter.cuh:20-26)


class PolySlice(ctypes.Structure):
    """common/src/typ.rs:13-40 / core/src/common/slice/poly_slice.h:3-13"""
    _fields_ = [("begin", ctypes.c_ulonglong), ("slice_len", ctypes.c_ulonglong),
                ("len", ctypes.c_ulonglong)]


class ImpNttItem(ctypes.Structure):
    """include/impeller_hip.h ImpNttItem"""
    _fields_ = [("data", ctypes.c_void_p), ("src", ctypes.c_void_p),
                ("log_n", ctypes.c_uint), ("log_src", ctypes.c_uint),
                ("twiddle", ctypes.c_void_p), ("inverse", ctypes.c_int),
                ("coset_shift", ctypes.c_void_p), ("plan", ctypes.c_uint * 10),
                ("plan_len", ctypes.c_uint)]


class ImpNttBatch(ctypes.Structure):
    _fields_ = [("dev", ctypes.c_void_p), ("count", ctypes.c_uint), ("nstages", ctypes.c_uint),
                ("total", ctypes.c_uint * 9), ("lds", ctypes.c_uint * 9)]


class ImpMsmItem(ctypes.Structure):
    """include/impeller_hip.h ImpMsmItem"""
    _fields_ = [("tables", ctypes.c_void_p), ("scalars", ctypes.c_void_p),
                ("len", ctypes.c_ulonglong), ("table_stride", ctypes.c_ulonglong),
                ("window", ctypes.c_uint), ("target_windows", ctypes.c_uint),
                ("task", ctypes.c_uint), ("result", ctypes.c_void_p)]


class ImpMsmBatch(ctypes.Structure):
    _fields_ = [("dev", ctypes.c_void_p), ("count", ctypes.c_uint), ("nstages", ctypes.c_uint),
                ("kind", ctypes.c_uint * 128), ("arg", ctypes.c_uint * 128),
                ("total", ctypes.c_uint * 128)]


_P = ctypes.c_void_p
_U32 = ctypes.c_uint
_U64 = ctypes.c_ulonglong
_INT = ctypes.c_int
_ERR = ctypes.c_int

# name -> (restype, argtypes); every symbol include/impeller_hip.h declares
SIGNATURES = {
    "ssip_ntt": (_ERR, [PolyPtr, _P, _U32, _P, _U32, _U32]),
    "ssip_precompute": (_ERR, [_P, _U32, _P]),
    "recompute_ntt": (_ERR, [PolyPtr, _P, _U32, _P, _U32, _P, _U32, _U32]),
    "gen_pq_omegas": (None, [_P, _P, _U32, _U32, _P]),
    "distribute_powers": (_ERR, [PolyPtr, _P, _U64, _P, ctypes.POINTER(PolySlice)]),
    "msm": (_ERR, [_P, _P, _U64, _U32, _U32, _U32, _U32, _U32, _P, _P, _U32, _P, _P]),
    "msm_precompute": (_ERR, [_U64, _P, _U32]),
    "msm_cfg": (_ERR, [_U32, _U32, _P, _P, _U64, _U32, _U32, _U32, _U32, _U32, _P, _P, _U32,
                       _P, _P]),
    "msm_precompute_cfg": (_ERR, [_U32, _U32, _U64, _P, _U32]),
    "fused_gate_abc_d_qe": (_ERR, [_P, _P, _U64, ctypes.POINTER(PolySlice), _P, _P]),
    "impeller_gate": (_ERR, [_P, _P, _U64, _P]),
    "poly_add": (_ERR, [PolyPtr, PolyPtr, PolyPtr, _P]),
    "poly_sub": (_ERR, [PolyPtr, PolyPtr, PolyPtr, _P]),
    "poly_mul": (_ERR, [PolyPtr, PolyPtr, PolyPtr, _P]),
    "poly_one_lagrange": (_ERR, [PolyPtr, _P]),
    "poly_one_coef": (_ERR, [PolyPtr, _P, ctypes.POINTER(PolySlice)]),
    "poly_zero": (_ERR, [PolyPtr, _P]),
    "poly_eval": (_ERR, [_P, ctypes.POINTER(ctypes.c_ulong), PolyPtr, _P, _P, _P]),
    "kate_division": (_ERR, [_P, ctypes.POINTER(ctypes.c_ulong), _U32, PolyPtr, _P, PolyPtr,
                             _P]),
    "scan_mul": (_ERR, [_P, ctypes.POINTER(ctypes.c_ulong), PolyPtr, _P]),
    "batched_invert": (_ERR, [_P, ctypes.POINTER(ctypes.c_ulong), PolyPtr, _P]),
    "inv_scalar": (_ERR, [_P, _P]),
    "permute": (_ERR, [_P, ctypes.POINTER(ctypes.c_ulong), ctypes.c_ulong, PolyPtr, PolyPtr,
                       PolyPtr, PolyPtr, _P]),
    "scalar_pow": (_ERR, [_P, _U64, _P]),
    "impeller_ntt": (_ERR, [_P, _P, _U32, _INT, _P, _P, _U32, _P]),
    "impeller_ntt_workspace_size": (ctypes.c_size_t, [_U32, _P, _U32]),
    "impeller_ntt_ws": (_ERR, [_P, _P, _U32, _INT, _P, _P, _U32, _P, ctypes.c_size_t, _P]),
    "impeller_ntt_noalloc": (_ERR, [_P, _P, _U32, _INT, _P, _P, _U32, _P]),
    "impeller_twiddles": (_ERR, [_P, _U64, _P, _P]),
    "impeller_ntt_batch_bytes": (ctypes.c_size_t, [_U32]),
    "impeller_ntt_batch_bytes_items": (ctypes.c_size_t, [ctypes.POINTER(ImpNttItem), _U32]),
    "impeller_ntt_batch_prepare": (_ERR, [ctypes.POINTER(ImpNttItem), _U32, _P, ctypes.c_size_t,
                                          ctypes.POINTER(ImpNttBatch), _P]),
    "impeller_ntt_batch_run": (_ERR, [ctypes.POINTER(ImpNttBatch), _U32, _P]),
    "impeller_graph_begin": (_ERR, [_P]),
    "impeller_graph_end": (_ERR, [_P, ctypes.POINTER(ctypes.c_void_p)]),
    "impeller_graph_launch": (_ERR, [_P, _P]),
    "impeller_graph_destroy": (_ERR, [_P]),
    "impeller_msm_workspace_size": (ctypes.c_size_t, [_U64, _U32]),
    "impeller_msm": (_ERR, [_P, _P, _U64, _U32, _P, _P, ctypes.c_size_t, _P]),
    "impeller_msm_cfg_workspace_size": (ctypes.c_size_t, [_U64, _U32, _U32, _U32]),
    "impeller_msm_tables": (_U32, [_U32, _U32]),
    "impeller_msm_precompute": (_ERR, [_P, _U64, _U32, _U32, _P]),
    "impeller_msm_cfg": (_ERR, [_P, _P, _U64, _U64, _U32, _U32, _U32, _P, _P, ctypes.c_size_t,
                                _P]),
    "impeller_msm_batch_bytes": (ctypes.c_size_t, [ctypes.POINTER(ImpMsmItem), _U32]),
    "impeller_msm_batch_prepare": (_ERR, [ctypes.POINTER(ImpMsmItem), _U32, _P, ctypes.c_size_t,
                                          ctypes.POINTER(ImpMsmBatch), _P]),
    "impeller_msm_batch_run": (_ERR, [ctypes.POINTER(ImpMsmBatch), _U32, _P]),
    "impeller_gen_scalars": (_ERR, [_P, _U64, _INT, _U64, _U64, _P]),
    "impeller_gen_bases": (_ERR, [_P, _U64, _U64, _U64, _P]),
    "impeller_field_op": (_ERR, [_INT, _INT, _P, _P, _P, _U64, _P]),
    "impeller_g1_op": (_ERR, [_INT, _P, _P, _P, _U64, _P]),
    "impeller_mont_mul_rate": (_ERR, [_INT, _U32, _U32, _P, _P]),
    "impeller_mlp_param_count": (_U32, [_U32, _U32]),
    "impeller_mlp_workspace_floats": (ctypes.c_size_t, [_U32, _U32, _U32]),
    "impeller_mlp_forward": (_ERR, [_P, _P, _U32, _U32, _U32, _P, _P, _P]),
    "impeller_mlp_grad": (_ERR, [_P, _P, _P, _U32, _U32, _U32, _P, _P, _P, _P]),
    "impeller_mlp_train_step": (_ERR, [_P, _P, _P, _P, _P, _U32, _U32, _U32, ctypes.c_float,
                                       _U32, _P, _P, _P]),
    "impeller_timing_enable": (None, [_INT]),
    "impeller_timing_reset": (None, []),
    "impeller_timing_read": (_INT, [_INT, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_ulonglong)]),
    "impeller_version": (ctypes.c_char_p, []),
    # device-runtime layer (csrc/runtime_api.hip)
    "impeller_rt_stream_create": (_ERR, [_INT, ctypes.POINTER(ctypes.c_void_p)]),
    "impeller_rt_stream_destroy": (_ERR, [_P]),
    "impeller_rt_stream_sync": (_ERR, [_P]),
    "impeller_rt_stream_alloc": (_ERR, [_P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "impeller_rt_stream_free": (_ERR, [_P, _P]),
    "impeller_rt_memcpy_async": (_ERR, [_INT, _P, _P, ctypes.c_size_t, _INT, _P]),
    "impeller_rt_event_create": (_ERR, [_INT, _INT, ctypes.POINTER(ctypes.c_void_p)]),
    "impeller_rt_event_record": (_ERR, [_P, _P]),
    "impeller_rt_event_sync": (_ERR, [_P]),
    "impeller_rt_stream_wait_event": (_ERR, [_P, _P]),
    "impeller_rt_event_reset": (_ERR, [_P]),
    "impeller_rt_event_elapsed": (_ERR, [_P, _P, ctypes.POINTER(ctypes.c_float)]),
    "impeller_rt_event_raw": (_P, [_P]),
    "impeller_rt_event_destroy": (_ERR, [_P]),
    "impeller_rt_alloc_pinned": (_ERR, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "impeller_rt_free_pinned": (_ERR, [_P]),
    "impeller_rt_static_create": (_ERR, [_INT, ctypes.c_size_t, _INT,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "impeller_rt_static_allocate": (_ERR, [_P, ctypes.c_size_t, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "impeller_rt_static_free": (_ERR, [_P, _P]),
    "impeller_rt_static_base": (_P, [_P]),
    "impeller_rt_static_destroy": (_ERR, [_P]),
    "impeller_rt_page_granularity": (_ERR, [_INT, ctypes.POINTER(ctypes.c_size_t)]),
    "impeller_rt_page_create": (_ERR, [_INT, ctypes.c_size_t, ctypes.c_size_t, _INT,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "impeller_rt_page_allocate": (_ERR, [_P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                         ctypes.c_size_t, ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "impeller_rt_page_deallocate": (_ERR, [_P, _P, ctypes.c_size_t]),
    "impeller_rt_page_size": (ctypes.c_size_t, [_P]),
    "impeller_rt_page_destroy": (_ERR, [_P]),
    "impeller_rt_num_cus": (_ERR, [_INT, ctypes.POINTER(ctypes.c_int)]),
    "impeller_rt_arch_name": (_ERR, [_INT, ctypes.c_char_p, ctypes.c_size_t]),
    "impeller_rt_file_read_to_device": (_ERR, [_INT, _U64, ctypes.c_size_t, _P, _P,
                                               ctypes.c_size_t, _P]),
    "impeller_rt_file_write_from_device": (_ERR, [_INT, _U64, ctypes.c_size_t, _P, _P,
                                                  ctypes.c_size_t, _P]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m impeller_amd.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(name, code):
    if code != 0:
        raise HipError(name, code)


def call(name, *args):
    fn = getattr(lib(), name)
    r = fn(*args)
    if SIGNATURES[name][0] is _ERR:
        check(name, r)
    return r


KERNEL_IDS = {"ntt_bitrev": 0, "ntt_dit": 1, "msm_digits": 2, "msm_bounds": 3,
              "msm_buckets": 5, "msm_reduce": 6, "msm_sum": 7, "msm_final": 8, "gate": 9,
              "msm_sort": 10, "msm_combine": 11}


def timing_enable(on=True):
    lib().impeller_timing_enable(1 if on else 0)


def timing_reset():
    lib().impeller_timing_reset()


def timing_read(name):
    """(total_ms, launches) for a kernel id since the last reset."""
    t = ctypes.c_double()
    c = ctypes.c_ulonglong()
    check("impeller_timing_read", lib().impeller_timing_read(KERNEL_IDS[name], ctypes.byref(t),
                                                             ctypes.byref(c)))
    return t.value, c.value
