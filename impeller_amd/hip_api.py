"""The device-runtime layer (impeller_rt_* in csrc/runtime_api.hip) with the
types of the reference's zkpoly_cuda_api crate: streams and events
(cuda_api/src/stream.rs), pinned host memory (mem.rs), the static and page
allocators (mem/static_allocator.rs, mem/page_allocator.rs), device facts
(device_info.rs) and file <-> device transfers (file.rs, cuFile there,
POSIX reads and writes here).  A failed call raises HipError, where the
reference's check macros panic."""
import ctypes
import os

from . import _lib

H2D, D2H, D2D = 1, 2, 3


def _out():
    return ctypes.c_void_p()


class HipStream:
    """CudaStream (stream.rs:137-257)."""

    def __init__(self, device: int = 0):
        h = _out()
        _lib.call("impeller_rt_stream_create", device, ctypes.byref(h))
        self.stream, self.device = h.value, device

    def raw(self):
        return self.stream

    def get_device(self):
        return self.device

    def sync(self):
        _lib.call("impeller_rt_stream_sync", self.stream)

    def allocate(self, nbytes: int) -> int:
        p = _out()
        _lib.call("impeller_rt_stream_alloc", self.stream, nbytes, ctypes.byref(p))
        return p.value

    def free(self, ptr: int):
        _lib.call("impeller_rt_stream_free", self.stream, ptr)

    def record(self, event: "HipEvent"):
        event.record(self)

    def wait(self, event: "HipEvent"):
        _lib.call("impeller_rt_stream_wait_event", self.stream, event.handle)

    wait_raw = wait

    def _copy(self, dst, src, nbytes, kind):
        _lib.call("impeller_rt_memcpy_async", self.device, dst, src, nbytes, kind, self.stream)

    def memcpy_h2d(self, dst, src, nbytes):
        self._copy(dst, src, nbytes, H2D)

    def memcpy_d2h(self, dst, src, nbytes):
        self._copy(dst, src, nbytes, D2H)

    def memcpy_d2d(self, dst, src, nbytes):
        self._copy(dst, src, nbytes, D2D)

    def destroy(self):
        if self.stream is not None:
            _lib.call("impeller_rt_stream_destroy", self.stream)
            self.stream = None


class HipEvent:
    """CudaEvent (host_ordered=True: waits block until it is recorded) or
    CudaEventRaw (host_ordered=False), stream.rs:1-135."""

    def __init__(self, device: int = 0, host_ordered: bool = True):
        h = _out()
        _lib.call("impeller_rt_event_create", device, 1 if host_ordered else 0, ctypes.byref(h))
        self.handle, self.device = h.value, device

    def get_device(self):
        return self.device

    def record(self, stream: HipStream):
        _lib.call("impeller_rt_event_record", self.handle, stream.raw())

    def sync(self):
        _lib.call("impeller_rt_event_sync", self.handle)

    def reset(self):
        _lib.call("impeller_rt_event_reset", self.handle)

    def elapsed(self, other: "HipEvent") -> float:
        ms = ctypes.c_float()
        _lib.call("impeller_rt_event_elapsed", self.handle, other.handle, ctypes.byref(ms))
        return ms.value

    def as_ptr(self):
        return _lib.lib().impeller_rt_event_raw(self.handle)

    def destroy(self):
        if self.handle is not None:
            _lib.call("impeller_rt_event_destroy", self.handle)
            self.handle = None


def alloc_pinned(nbytes: int) -> int:
    """mem.rs:18-33."""
    p = _out()
    _lib.call("impeller_rt_alloc_pinned", nbytes, ctypes.byref(p))
    return p.value


def free_pinned(ptr: int):
    _lib.call("impeller_rt_free_pinned", ptr)


class StaticAllocator:
    """mem/static_allocator.rs: offsets chosen by the caller's memory plan."""

    def __init__(self, device: int, max_size: int, check_overlap: bool):
        h = _out()
        _lib.call("impeller_rt_static_create", device, max_size, 1 if check_overlap else 0,
                  ctypes.byref(h))
        self.handle, self.max_size = h.value, max_size

    def base(self) -> int:
        return _lib.lib().impeller_rt_static_base(self.handle)

    def allocate(self, offset: int, nbytes: int) -> int:
        p = _out()
        _lib.call("impeller_rt_static_allocate", self.handle, offset, nbytes, ctypes.byref(p))
        return p.value

    def free(self, ptr: int):
        _lib.call("impeller_rt_static_free", self.handle, ptr)

    def destroy(self):
        if self.handle is not None:
            _lib.call("impeller_rt_static_destroy", self.handle)
            self.handle = None


def page_granularity(device: int = 0) -> int:
    g = ctypes.c_size_t()
    _lib.call("impeller_rt_page_granularity", device, ctypes.byref(g))
    return g.value


class PageAllocator:
    """mem/page_allocator.rs over HIP virtual memory (device pages)."""

    def __init__(self, device: int, page_size: int, page_num: int, memory_check: bool):
        h = _out()
        _lib.call("impeller_rt_page_create", device, page_size, page_num,
                  1 if memory_check else 0, ctypes.byref(h))
        self.handle = h.value

    def page_size(self) -> int:
        return _lib.lib().impeller_rt_page_size(self.handle)

    def allocate(self, va_size: int, page_ids, vaid: int) -> int:
        ids = (ctypes.c_size_t * max(1, len(page_ids)))(*page_ids)
        p = _out()
        _lib.call("impeller_rt_page_allocate", self.handle, va_size, ids, len(page_ids), vaid,
                  ctypes.byref(p))
        return p.value

    def deallocate(self, ptr: int, vaid: int):
        _lib.call("impeller_rt_page_deallocate", self.handle, ptr, vaid)

    def destroy(self):
        if self.handle is not None:
            _lib.call("impeller_rt_page_destroy", self.handle)
            self.handle = None


def get_num_cus(device: int = 0) -> int:
    """device_info.rs get_num_sms."""
    n = ctypes.c_int()
    _lib.call("impeller_rt_num_cus", device, ctypes.byref(n))
    return n.value


def arch_name(device: int = 0) -> str:
    """device_info.rs sm_version: the gfx target."""
    buf = ctypes.create_string_buffer(64)
    _lib.call("impeller_rt_arch_name", device, buf, len(buf))
    return buf.value.decode()


def read_file_to_device(fd: int, file_offset: int, nbytes: int, dst: int, staging: int,
                        staging_bytes: int, stream: HipStream):
    """file.rs (cuFile read) through a pinned staging buffer."""
    _lib.call("impeller_rt_file_read_to_device", fd, file_offset, nbytes, dst, staging,
              staging_bytes, stream.raw())


def write_file_from_device(fd: int, file_offset: int, nbytes: int, src: int, staging: int,
                           staging_bytes: int, stream: HipStream):
    _lib.call("impeller_rt_file_write_from_device", fd, file_offset, nbytes, src, staging,
              staging_bytes, stream.raw())


def open_file(path: str, write: bool = False) -> int:
    """file.rs register_cufile: here just the descriptor the transfers use."""
    return os.open(path, (os.O_RDWR | os.O_CREAT) if write else os.O_RDONLY, 0o644)
