"""Polynomial operators, mirroring the reference's Rust wrappers in
core/src/poly.rs (PolyAdd, PolySub, PolyMul, PolyZero, PolyOneLagrange,
PolyOneCoef, PolyEval, KateDivision, PolyScan, PolyInvert, PolyPermute,
ScalarInv, ScalarPow).

Each class exposes get_fn() returning a callable with the reference
runtime's Function signature f(mut_var, var, gpu_mapping, poly_slice)
(runtime/src/functions.rs:20-36) over the same argument lists, and the
get_buffer_size() queries of the ones that need a temporary buffer.  Every
call goes through the C ABI (libimpeller_hip.so).  Only device-resident
variables are accepted: the reference's host-side branches for host
variables belong to its CPU runtime, which this library does not replace.
"""
import ctypes

from . import _lib
from .runtime import copy_elements, copy_view


def _gpu(*vs):
    for v in vs:
        if getattr(v, "device", "cuda") == "cpu":
            raise NotImplementedError("host-resident variables are not handled by this library")


def _size_query(name, *args):
    size = ctypes.c_ulong(0)
    _lib.call(name, None, ctypes.byref(size), *args)
    return size.value


class _Binary:
    """poly.rs:714-844: mut_var=[res], var=[a, b, stream]."""
    symbol = ""

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 1 and len(var) == 3
            res, a, b, stream = mut_var[0], var[0], var[1], var[2]
            _gpu(res, a, b)
            _lib.call(self.symbol, res.poly_ptr(), a.poly_ptr(), b.poly_ptr(), stream.raw())
        return f


class PolyAdd(_Binary):
    symbol = "poly_add"


class PolySub(_Binary):
    symbol = "poly_sub"


class PolyMul(_Binary):
    symbol = "poly_mul"


class PolyZero:
    """poly.rs:626-657: mut_var=[target], var=[stream]."""

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 1 and len(var) == 1
            _gpu(mut_var[0])
            _lib.call("poly_zero", mut_var[0].poly_ptr(), var[0].raw())
        return f


class PolyOneLagrange:
    """poly.rs:532-561: mut_var=[target], var=[stream]."""

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 1 and len(var) == 1
            _gpu(mut_var[0])
            _lib.call("poly_one_lagrange", mut_var[0].poly_ptr(), var[0].raw())
        return f


class PolyOneCoef:
    """poly.rs:563-624: one at the slice's origin (to_origin(i) == 0), zero
    elsewhere; mut_var=[target], var=[stream]."""

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 1 and len(var) == 1
            target = mut_var[0]
            _gpu(target)
            if poly_slice is not None:
                assert poly_slice.slice_len == target.len
            else:
                poly_slice = _lib.PolySlice(0, target.len, target.len)
            _lib.call("poly_one_coef", target.poly_ptr(), var[0].raw(), ctypes.byref(poly_slice))
        return f


class PolyEval:
    """poly.rs:667-712: mut_var=[temp_buf, res], var=[poly, x, stream]."""

    def get_buffer_size(self, length):
        return _size_query("poly_eval", _lib.PolyPtr(None, length, 0, 0, length), None, None,
                           None)

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 2 and len(var) == 3
            temp, res = mut_var
            poly, x, stream = var
            _gpu(poly)
            _lib.call("poly_eval", temp.data_ptr(), None, poly.poly_ptr(), res.data_ptr(),
                      x.data_ptr(), stream.raw())
        return f


class KateDivision:
    """poly.rs:659-665,714-760 region: mut_var=[temp_buf, res], var=[p, b,
    stream]; p.len a power of two, res.len == p.len."""

    def get_buffer_size(self, log_p):
        n = 1 << log_p
        return _size_query("kate_division", log_p, _lib.PolyPtr(None, n, 0, 0, n), None,
                           _lib.PolyPtr(None, n, 0, 0, n), None)

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 2 and len(var) == 3
            temp, res = mut_var
            p, b, stream = var
            _gpu(res, p)
            assert res.len == p.len
            assert p.len & (p.len - 1) == 0 and p.len > 0
            log_p = p.len.bit_length() - 1
            _lib.call("kate_division", temp.data_ptr(), None, log_p, p.poly_ptr(),
                      b.data_ptr(), res.poly_ptr(), stream.raw())
        return f


class PolyScan:
    """poly.rs:423-480: mut_var=[temp_buf, target], var=[p, x0, stream]:
    target[0] = x0, target[i] = x0 * p[0] * ... * p[i-1]."""

    def get_buffer_size(self, length):
        return _size_query("scan_mul", _lib.PolyPtr(None, length, 0, 0, length), None)

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 2 and len(var) == 3
            temp, target = mut_var
            p, x0, stream = var
            _gpu(target, p)
            assert target.len == p.len
            # target[0] = x0; target[1..] = p[0..len-1) (poly.rs:457-462)
            copy_elements(x0.values, target, 0, stream)
            copy_view(p, 0, p.len - 1, target, 1, stream)
            _lib.call("scan_mul", temp.data_ptr(), None, target.poly_ptr(), stream.raw())
        return f


class PolyInvert:
    """poly.rs:380-421: mut_var=[temp_buf, target], var=[stream]; non-zero
    elements inverted in place, zeros kept."""

    def get_buffer_size(self, length):
        return _size_query("batched_invert", _lib.PolyPtr(None, length, 0, 0, length), None)

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 2 and len(var) == 1
            temp, target = mut_var
            _gpu(target)
            _lib.call("batched_invert", temp.data_ptr(), None, target.poly_ptr(), var[0].raw())
        return f


class PolyPermute:
    """poly.rs:227-282: mut_var=[res_input, res_table, temp_buf],
    var=[input, table, stream]; usable = res_input.len."""

    def get_buffer_size(self, usable_len):
        v = _lib.PolyPtr(None, usable_len, 0, 0, usable_len)
        return _size_query("permute", usable_len, v, v, v, v, None)

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 3 and len(var) == 3
            res_input, res_table, temp = mut_var
            inp, table, stream = var
            _gpu(res_input, res_table, inp, table)
            _lib.call("permute", temp.data_ptr(), None, res_input.len, inp.poly_ptr(),
                      table.poly_ptr(), res_input.poly_ptr(), res_table.poly_ptr(), stream.raw())
        return f


class ScalarInv:
    """poly.rs:343-378: mut_var=[scalar], var=[stream]."""

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 1 and len(var) == 1
            _lib.call("inv_scalar", mut_var[0].data_ptr(), var[0].raw())
        return f


class ScalarPow:
    """poly.rs:284-341: mut_var=[scalar], var=[stream]; scalar := scalar^exp."""

    def __init__(self, exp):
        self.exp = int(exp) & ((1 << 64) - 1)

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            assert len(mut_var) == 1 and len(var) == 1
            _lib.call("scalar_pow", mut_var[0].data_ptr(), self.exp, var[0].raw())
        return f
