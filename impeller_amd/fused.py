"""Fused element-wise gate, mirroring core/src/fused_kernels.rs:95-200.

The reference closure copies ConstPolyPtr/PolyPtr arrays into a device
arg_buffer and calls name(vars_dev, mut_vars_dev, len, &slice,
local_buffer, stream); argument convention arg = [stream, inputs...],
arg_mut = [arg_buffer, local_buffer, outputs...] (emit_func.rs:36-56).
"""
import ctypes

import numpy as np

from . import _lib


class FusedGate:
    """out = a*b*c - d + q*e*(1 - q)  (symbol fused_gate_abc_d_qe)."""
    name = "fused_gate_abc_d_qe"
    n_inputs = 6
    n_outputs = 1

    def get_fn(self):
        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            import torch
            stream = var[0]
            inputs = var[1:]
            arg_buffer, local_buffer, outputs = mut_var[0], mut_var[1], mut_var[2:]
            assert len(inputs) == self.n_inputs and len(outputs) == self.n_outputs
            length = outputs[0].len
            ptrs = [x.poly_ptr() for x in inputs] + [y.poly_ptr() for y in outputs]
            host = (_lib.PolyPtr * len(ptrs))(*ptrs)
            raw = np.frombuffer(bytes(host), dtype=np.uint8)
            nbytes = raw.nbytes
            assert arg_buffer.numel() * arg_buffer.element_size() >= nbytes
            arg_buffer.view(torch.uint8)[:nbytes].copy_(torch.from_numpy(raw.copy()))
            base = arg_buffer.data_ptr()
            sl = poly_slice or _lib.PolySlice(0, length, length)
            _lib.call(self.name, base, base + self.n_inputs * ctypes.sizeof(_lib.PolyPtr),
                      length, ctypes.byref(sl),
                      local_buffer.data_ptr() if local_buffer is not None else None,
                      stream.raw())
        return f


def gate_device(inputs, out, length, stream):
    """out = a*b*c - d + q*e*(1 - q) over contiguous device tensors
    (impeller_gate; kernel launch only, graph-capturable)."""
    ptrs = (ctypes.c_void_p * 6)(*[t.data_ptr() for t in inputs])
    _lib.call("impeller_gate", ptrs, out.data_ptr(), length, stream.raw())
