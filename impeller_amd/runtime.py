"""Host-side mirror of the reference runtime's array and stream types.

ScalarArray follows runtime/src/scalar.rs:182-191,400-477 (len, rotate,
slice_info {offset, whole_len}; rotate only on unsliced arrays; slice of a
rotated array records the physical start), PointArray runtime/src/point.rs:
58-63, and Stream the CudaStream wrapper of cuda_api/src/stream.rs (here a
torch HIP stream).  Device memory is a torch tensor on the current HIP
device; torch is plumbing only, every computation goes through the C ABI.
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib

FIELD_LIMBS = 8
POINT_LIMBS = 16


def _torch():
    import torch
    return torch


@dataclass
class ScalarSlice:
    offset: int
    whole_len: int


class Stream:
    """A HIP stream (cuda_api/src/stream.rs CudaStream)."""

    def __init__(self, device: int = 0, stream=None):
        torch = _torch()
        self.device = device
        self.torch_stream = stream if stream is not None else torch.cuda.Stream(device=device)

    @classmethod
    def current(cls, device: int = 0):
        torch = _torch()
        return cls(device, torch.cuda.current_stream(device))

    def raw(self):
        return self.torch_stream.cuda_stream

    def get_device(self):
        return self.device

    def sync(self):
        self.torch_stream.synchronize()

    def allocate(self, n_limbs: int):
        torch = _torch()
        return torch.empty(n_limbs, dtype=torch.int32, device=f"cuda:{self.device}")


class ScalarArray:
    """An array of n Fr elements (32 B each), on the host (numpy) or device (torch)."""

    def __init__(self, length: int, values=None, device: str = "cpu", device_id: int = 0):
        self.len = length
        self.rotate_ = 0
        self.slice_info: Optional[ScalarSlice] = None
        self.device = device
        if values is None:
            if device == "cpu":
                values = np.zeros((length, FIELD_LIMBS), dtype=np.uint32)
            else:
                torch = _torch()
                values = torch.zeros(length * FIELD_LIMBS, dtype=torch.int32,
                                     device=f"cuda:{device_id}")
        self.values = values

    # --- reference semantics (runtime/src/scalar.rs:396-477)
    def get_rotation(self):
        return self.rotate_

    def rotate(self, shift: int):
        assert self.slice_info is None
        self.rotate_ = (self.rotate_ + shift) % self.len

    def slice(self, start: int, end: int) -> "ScalarArray":
        assert start <= end <= self.len
        if self.slice_info is not None:
            raise ValueError("this is already a slice, can't be sliced again")
        actual_pos = (self.len + start - self.rotate_) % self.len
        s = ScalarArray.__new__(ScalarArray)
        s.len = end - start
        s.rotate_ = 0
        s.values = self.values
        s.device = self.device
        s.slice_info = ScalarSlice(actual_pos, self.len)
        return s

    # --- FFI view (core/src/poly_ptr.rs:47-66)
    def data_ptr(self):
        if self.device == "cpu":
            return self.values.ctypes.data
        return self.values.data_ptr()

    def poly_ptr(self) -> _lib.PolyPtr:
        if self.slice_info is None:
            offset, whole = 0, self.len
        else:
            offset, whole = self.slice_info.offset, self.slice_info.whole_len
        return _lib.PolyPtr(self.data_ptr(), self.len, self.rotate_, offset, whole)

    # --- transfers
    @classmethod
    def from_numpy(cls, arr, device: str = "cuda", device_id: int = 0):
        arr = np.ascontiguousarray(arr, dtype=np.uint32).reshape(-1, FIELD_LIMBS)
        if device == "cpu":
            return cls(arr.shape[0], arr.copy(), "cpu")
        torch = _torch()
        t = torch.from_numpy(arr.reshape(-1).view(np.int32).copy()).to(f"cuda:{device_id}")
        return cls(arr.shape[0], t, "cuda", device_id)

    def to_numpy(self):
        """The whole physical buffer as (whole_len, 8) uint32."""
        if self.device == "cpu":
            return np.asarray(self.values).reshape(-1, FIELD_LIMBS)
        return self.values.cpu().numpy().view(np.uint32).reshape(-1, FIELD_LIMBS)

    def logical_numpy(self):
        """Elements in logical order (rotation / slice applied)."""
        whole = self.to_numpy()
        if self.slice_info is None:
            wl, off = self.len, (self.len - self.rotate_) % self.len if self.rotate_ else 0
        else:
            wl, off = self.slice_info.whole_len, self.slice_info.offset
        idx = (off + np.arange(self.len)) % wl
        return whole[idx]


class PointArray:
    """n affine G1 points (64 B each) -- runtime/src/point.rs:58-63."""

    def __init__(self, length: int, values):
        self.len = length
        self.values = values

    @classmethod
    def from_numpy(cls, arr, device: str = "cuda", device_id: int = 0):
        arr = np.ascontiguousarray(arr, dtype=np.uint32).reshape(-1, POINT_LIMBS)
        if device == "cpu":
            return cls(arr.shape[0], arr.copy())
        torch = _torch()
        t = torch.from_numpy(arr.reshape(-1).view(np.int32).copy()).to(f"cuda:{device_id}")
        return cls(arr.shape[0], t)

    def data_ptr(self):
        if isinstance(self.values, np.ndarray):
            return self.values.ctypes.data
        return self.values.data_ptr()

    def to_numpy(self):
        if isinstance(self.values, np.ndarray):
            return self.values.reshape(-1, POINT_LIMBS)
        return self.values.cpu().numpy().view(np.uint32).reshape(-1, POINT_LIMBS)


def device_u32(n_words: int, device_id: int = 0):
    torch = _torch()
    return torch.empty(n_words, dtype=torch.int32, device=f"cuda:{device_id}")


def to_device(arr, device_id: int = 0):
    torch = _torch()
    a = np.ascontiguousarray(arr, dtype=np.uint32).reshape(-1)
    return torch.from_numpy(a.view(np.int32).copy()).to(f"cuda:{device_id}")


def to_host(t, cols=FIELD_LIMBS):
    return t.cpu().numpy().view(np.uint32).reshape(-1, cols)


class Scalar:
    """One Fr element on the device (runtime/src/scalar.rs:27-30 Scalar with
    device = GPU): `value` is a torch tensor of 8 words."""

    def __init__(self, values=None, device_id: int = 0):
        if values is None:
            torch = _torch()
            values = torch.zeros(FIELD_LIMBS, dtype=torch.int32, device=f"cuda:{device_id}")
        self.values = values
        self.device = "cuda"

    @classmethod
    def from_numpy(cls, arr, device_id: int = 0):
        torch = _torch()
        a = np.ascontiguousarray(arr, dtype=np.uint32).reshape(FIELD_LIMBS)
        return cls(torch.from_numpy(a.view(np.int32).copy()).to(f"cuda:{device_id}"))

    def data_ptr(self):
        return self.values.data_ptr()

    def to_numpy(self):
        return self.values.cpu().numpy().view(np.uint32).reshape(FIELD_LIMBS)


class GpuBuffer:
    """A device byte buffer (runtime/src/gpu_buffer.rs GpuBuffer)."""

    def __init__(self, size: int, device_id: int = 0):
        torch = _torch()
        self.size = size
        # 256-byte aligned: a torch allocation of at least one word
        self.values = torch.empty(max(size, 1), dtype=torch.uint8, device=f"cuda:{device_id}")

    def data_ptr(self):
        return self.values.data_ptr()


def _phys_segments(arr: "ScalarArray", start: int, count: int):
    """Physical (first, length) pieces holding logical elements
    [start, start + count) of a view (iter.cuh:309-317 index map)."""
    if arr.slice_info is None:
        wl = arr.len
        off = (arr.len - arr.rotate_) % arr.len if arr.rotate_ else 0
    else:
        wl, off = arr.slice_info.whole_len, arr.slice_info.offset
    pieces, p = [], (off + start) % wl if wl else 0
    while count > 0:
        k = min(count, wl - p)
        pieces.append((p, k))
        count -= k
        p = 0
    return pieces


def copy_view(src: "ScalarArray", s0: int, count: int, dst: "ScalarArray", d0: int,
              stream: "Stream"):
    """Device-to-device copy of logical elements src[s0 : s0 + count] to
    dst[d0 : d0 + count] on stream (ScalarArray::gpu2gpu of slices,
    runtime/src/scalar.rs), split where either view wraps."""
    torch = _torch()
    a, b = _phys_segments(src, s0, count), _phys_segments(dst, d0, count)
    with torch.cuda.stream(stream.torch_stream):
        ia = ib = 0
        pa, ka = a[0] if a else (0, 0)
        pb, kb = b[0] if b else (0, 0)
        while count > 0:
            k = min(ka, kb)
            dst.values[8 * pb: 8 * (pb + k)].copy_(src.values[8 * pa: 8 * (pa + k)],
                                                   non_blocking=True)
            count -= k
            pa, ka, pb, kb = pa + k, ka - k, pb + k, kb - k
            if ka == 0 and count > 0:
                ia += 1
                pa, ka = a[ia]
            if kb == 0 and count > 0:
                ib += 1
                pb, kb = b[ib]


def copy_elements(values, dst: "ScalarArray", d0: int, stream: "Stream"):
    """Copies a device tensor of whole elements into logical dst[d0 : ...]."""
    torch = _torch()
    n = values.numel() // FIELD_LIMBS
    with torch.cuda.stream(stream.torch_stream):
        done = 0
        for p, k in _phys_segments(dst, d0, n):
            dst.values[8 * p: 8 * (p + k)].copy_(values[8 * done: 8 * (done + k)],
                                                 non_blocking=True)
            done += k
