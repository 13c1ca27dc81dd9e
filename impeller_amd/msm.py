"""MSM operators, mirroring core/src/msm.rs and common/src/msm_config.rs.

MSM.get_fn() keeps the reference runtime's argument convention
(compiler/src/transit/type3/lowering/emit_func.rs:310-337):
  var     = [point tables x n_precompute (host PointArray), scalar arrays x batches]
  mut_var = [device buffer x cards, result point x batches]
and the buffer-size query protocol (core/src/msm.rs:282-302).  The C symbol
is msm_cfg with the configuration passed explicitly instead of baked into a
per-configuration .so name.
"""
import ctypes
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _lib


@dataclass
class MsmConfig:
    """common/src/msm_config.rs:4-56"""
    window_size: int = 16
    target_window: int = 16
    cards: List[int] = field(default_factory=lambda: [0])
    debug: bool = False
    batch_per_run: int = 1
    parts: int = 1
    stage_scalars: int = 1
    stage_points: int = 1
    bits: int = 254

    def actual_windows(self):
        # the device kernels use floor(bits / c) + 1 signed windows so the
        # top digit never carries (ceil in the reference plus its assert)
        return self.bits // self.window_size + 1

    def n_windows(self):
        a = self.actual_windows()
        return a if a < self.target_window else self.target_window

    def get_precompute(self):
        a = self.actual_windows()
        return -(-a // self.n_windows())


def _ptr_array(ptrs):
    arr = (ctypes.c_void_p * max(len(ptrs), 1))(*ptrs)
    return arr


class MSM:
    def __init__(self, config: MsmConfig):
        self.config = config

    def get_buffer_size(self, length: int):
        c = self.config
        sizes = (ctypes.c_ulong * len(c.cards))()
        cards = (ctypes.c_uint * len(c.cards))(*c.cards)
        _lib.call("msm_cfg", c.window_size, c.target_window, None, sizes, length,
                  c.batch_per_run, c.parts, c.stage_scalars, c.stage_points, len(c.cards),
                  cards, None, 0, None, None)
        return list(sizes)

    def get_fn(self):
        c = self.config
        n_pre = c.get_precompute()

        def f(mut_var, var, gpu_mapping=lambda x: x, poly_slice=None):
            batches = len(var) - n_pre
            assert len(mut_var) == batches + len(c.cards)
            length = var[0].len
            h_points = [var[i].data_ptr() for i in range(n_pre)]
            h_scalars = []
            for i in range(n_pre, len(var)):
                assert var[i].get_rotation() == 0, "currently, we don't support rotate in msm"
                h_scalars.append(var[i].data_ptr())
            buffers = [mut_var[i].data_ptr() for i in range(len(c.cards))]
            results = [mut_var[len(c.cards) + b].ctypes.data for b in range(batches)]
            cards = (ctypes.c_uint * len(c.cards))(*[gpu_mapping(x) for x in c.cards])
            _lib.call("msm_cfg", c.window_size, c.target_window, _ptr_array(buffers), None,
                      length, c.batch_per_run, c.parts, c.stage_scalars, c.stage_points,
                      len(c.cards), cards, _ptr_array(h_points), batches, _ptr_array(h_scalars),
                      _ptr_array(results))
        return f


class MSMPrecompute:
    """core/src/msm.rs:137-191: fills point tables 1.. from table 0 (host)."""

    def __init__(self, config: MsmConfig):
        self.config = config

    def get_fn(self):
        c = self.config

        def f(targets, length, max_cards=1):
            ptrs = [t.data_ptr() for t in targets]
            assert len(ptrs) == c.get_precompute()
            _lib.call("msm_precompute_cfg", c.window_size, c.target_window, length,
                      _ptr_array(ptrs), max_cards)
        return f


# ------------------------------------------------------ device-resident API
def workspace_size(n: int, window: int) -> int:
    return int(_lib.call("impeller_msm_workspace_size", n, window))


def cfg_workspace_size(n: int, window: int, target: int = 0, task: int = 64) -> int:
    return int(_lib.call("impeller_msm_cfg_workspace_size", n, window, target, task))


def n_tables(window: int, target: int) -> int:
    return int(_lib.call("impeller_msm_tables", window, target))


def precompute_device(tables_t, n, window, target, stream):
    """Fill tables 1.. of a device tensor holding n_tables * n affine points."""
    _lib.call("impeller_msm_precompute", tables_t.data_ptr(), n, window, target, stream.raw())


def msm_device_cfg(tables_t, scalars_t, n, window, target, stream, task=64, workspace=None,
                   out=None, sync=True, table_stride=0):
    """Device MSM over precomputed tables; returns the (16,) affine result."""
    from .runtime import device_u32, to_host
    ws_bytes = cfg_workspace_size(n, window, target, task)
    if workspace is None:
        workspace = device_u32((ws_bytes + 3) // 4)
    if out is None:
        out = device_u32(16)
    _lib.call("impeller_msm_cfg", tables_t.data_ptr(), scalars_t.data_ptr(), n, table_stride,
              window, target, task, out.data_ptr(), workspace.data_ptr(), ws_bytes, stream.raw())
    if not sync:
        return out
    stream.sync()
    return to_host(out, 16)[0]


def msm_device(bases_t, scalars_t, n, window, stream, workspace=None):
    """Device MSM; returns the affine result as a (16,) uint32 numpy array."""
    from .runtime import device_u32, to_host
    ws_bytes = workspace_size(n, window)
    if workspace is None:
        workspace = device_u32((ws_bytes + 3) // 4)
    out = device_u32(16)
    _lib.call("impeller_msm", bases_t.data_ptr(), scalars_t.data_ptr(), n, window,
              out.data_ptr(), workspace.data_ptr(), ws_bytes, stream.raw())
    stream.sync()
    return to_host(out, 16)[0]


class MsmBatch:
    """A prepared batch of MSMs (impeller_msm_batch_*).

    items: dicts with tables (device tensor), scalars (device tensor), len,
    window, target (0: per-window buckets, 1: precomputed tables), task,
    table_stride (points between tables of a shared set; default len).
    Results land in self.results (device, 16 words per MSM).
    """

    def __init__(self, items, stream):
        from .runtime import device_u32
        n = len(items)
        self.results = device_u32(16 * n)
        arr = (_lib.ImpMsmItem * n)()
        self._keep = list(items)
        for i, it in enumerate(items):
            a = arr[i]
            a.tables = it["tables"].data_ptr()
            a.scalars = it["scalars"].data_ptr()
            a.len = it["len"]
            a.table_stride = it.get("table_stride", 0)
            a.window = it["window"]
            a.target_windows = it.get("target", 0)
            a.task = it.get("task", 64)
            a.result = self.results.data_ptr() + 64 * i
        nbytes = int(_lib.call("impeller_msm_batch_bytes", arr, n))
        if nbytes == 0:
            raise ValueError("invalid MSM batch item")
        self.dev = device_u32((nbytes + 3) // 4)
        self.info = _lib.ImpMsmBatch()
        _lib.call("impeller_msm_batch_prepare", arr, n, self.dev.data_ptr(), nbytes,
                  ctypes.byref(self.info), stream.raw())

    def run(self, stream, max_blocks=0):
        _lib.call("impeller_msm_batch_run", ctypes.byref(self.info), max_blocks, stream.raw())

    def results_host(self):
        from .runtime import to_host
        return to_host(self.results, 16)
