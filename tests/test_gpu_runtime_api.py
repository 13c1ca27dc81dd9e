"""The device-runtime layer (impeller_rt_*) on the GPU, following the
reference crate's own tests: cuda_api/src/stream.rs:267-277 (stream, event,
stream-ordered allocation), mem/static_allocator.rs:91-158 (round trip,
overlap rules), mem/page_allocator.rs:379-560 (write / read through mapped
pages, address reuse per vaid, unaligned page size, short reservation), plus
host-ordered event waits across threads, kernels writing into mapped pages,
device facts and file <-> device transfers."""
import ctypes
import threading
import time

import numpy as np
import pytest
import torch

from impeller_amd import _lib
from impeller_amd import hip_api as H

pytestmark = pytest.mark.gpu


def h2d_d2h(stream, dptr, data):
    stream.memcpy_h2d(dptr, data.ctypes.data, data.nbytes)
    back = np.zeros_like(data)
    stream.memcpy_d2h(back.ctypes.data, dptr, data.nbytes)
    stream.sync()
    return back


def test_stream_event_alloc(stream):
    # stream.rs:267-277
    s = H.HipStream(0)
    e = H.HipEvent(0)
    s.record(e)
    s.wait(e)
    s.sync()
    p = s.allocate(1 << 30)
    data = np.arange(1 << 20, dtype=np.uint32)
    assert (h2d_d2h(s, p, data) == data).all()
    q = s.allocate(data.nbytes)
    s.memcpy_d2d(q, p, data.nbytes)
    back = np.zeros_like(data)
    s.memcpy_d2h(back.ctypes.data, q, data.nbytes)
    s.sync()
    assert (back == data).all()
    s.free(q)
    s.free(p)
    s.sync()
    e.destroy()
    s.destroy()


def test_host_ordered_event_waits_for_record():
    s = H.HipStream(0)
    e = H.HipEvent(0, host_ordered=True)
    done = []

    def waiter():
        e.sync()          # blocks until another thread records e
        done.append(time.time())

    t = threading.Thread(target=waiter, daemon=True)
    t.start()
    time.sleep(0.3)
    assert t.is_alive() and not done   # not recorded yet: still waiting
    t_rec = time.time()
    e.record(s)
    t.join(timeout=30)
    assert not t.is_alive() and done and done[0] >= t_rec
    # after reset a new wait needs a new record; a second stream waits on it
    e.reset()
    s2 = H.HipStream(0)
    e.record(s)
    s2.wait(e)
    s2.sync()
    # raw events: elapsed time between two records around a copy
    a, b = H.HipEvent(0, host_ordered=False), H.HipEvent(0, host_ordered=False)
    p = s.allocate(64 << 20)
    q = s.allocate(64 << 20)
    a.record(s)
    s.memcpy_d2d(q, p, 64 << 20)
    b.record(s)
    b.sync()
    assert a.elapsed(b) >= 0.0
    s.free(p)
    s.free(q)
    s.sync()
    for x in (a, b, e):
        x.destroy()
    s2.destroy()
    s.destroy()


def test_pinned_memory_round_trip():
    s = H.HipStream(0)
    n = 3 << 20
    hp = H.alloc_pinned(n)
    ctypes.memmove(hp, bytes(range(256)) * (n // 256), n)
    d = s.allocate(n)
    s.memcpy_h2d(d, hp, n)
    back = np.zeros(n, np.uint8)
    s.memcpy_d2h(back.ctypes.data, d, n)
    s.sync()
    assert (back == np.tile(np.arange(256, dtype=np.uint8), n // 256)).all()
    s.free(d)
    s.sync()
    H.free_pinned(hp)
    s.destroy()


def test_static_allocator():
    s = H.HipStream(0)
    # static_allocator.rs:91-106: round trip at offset 0
    a = H.StaticAllocator(0, 1 << 20, False)
    p = a.allocate(0, 9 * 4)
    va = np.arange(1, 10, dtype=np.int32)
    assert (h2d_d2h(s, p, va) == va).all()
    a.free(p)
    assert a.allocate(4096, 16) == a.base() + 4096
    a.destroy()
    # static_allocator.rs:108-138: adjacent ranges, reuse after free
    a = H.StaticAllocator(0, 1024, True)
    p1, p2 = a.allocate(0, 16), a.allocate(16, 16)
    assert p2 == p1 + 16
    a.free(p1)
    p3 = a.allocate(0, 16)
    a.free(p2)
    a.free(p3)
    p4, p5 = a.allocate(0, 16), a.allocate(16, 16)
    # static_allocator.rs:140-158: an overlapping range is refused
    with pytest.raises(_lib.HipError):
        a.allocate(8, 16)
    with pytest.raises(_lib.HipError):
        a.allocate(1000, 100)          # past the arena
    with pytest.raises(_lib.HipError):
        a.free(p1 + 4)                 # no allocation starts there
    a.free(p4)
    a.free(p5)
    a.destroy()
    s.destroy()


def page_size_for(g):
    return max(g, 2 << 20) // g * g


def test_page_allocator_maps_pages():
    s = H.HipStream(0)
    g = H.page_granularity(0)
    assert g > 0
    ps = page_size_for(g)
    a = H.PageAllocator(0, ps, 10, True)
    assert a.page_size() == ps
    # page_allocator.rs:445-461: two pages in a reservation twice as large
    p = a.allocate(4 * ps, [0, 1], 0)
    data = (np.arange(2 * ps, dtype=np.uint64) % 256).astype(np.uint8)
    assert (h2d_d2h(s, p, data) == data).all()
    # a kernel of this library writes into the mapped pages
    n = ps // 32
    _lib.call("impeller_gen_scalars", p, n, 0, 99, 0, s.raw())
    ref = torch.empty(n * 8, dtype=torch.int32, device="cuda")
    _lib.call("impeller_gen_scalars", ref.data_ptr(), n, 0, 99, 0,
              torch.cuda.current_stream().cuda_stream)
    got = np.zeros(n * 8, np.uint32)
    s.memcpy_d2h(got.ctypes.data, p, n * 32)
    s.sync()
    torch.cuda.synchronize()
    assert (got == ref.cpu().numpy().view(np.uint32)).all()
    # a page is one address's at a time
    with pytest.raises(_lib.HipError):
        a.allocate(2 * ps, [1, 2], 1)
    a.deallocate(p, 0)
    # page_allocator.rs:463-478: the same vaid gives back the same address
    # (and must name the pages mapped there)
    with pytest.raises(_lib.HipError):
        a.allocate(4 * ps, [1, 0], 0)
    p2 = a.allocate(4 * ps, [0, 1], 0)
    assert p2 == p
    assert (h2d_d2h(s, p2, data) == data).all()
    # destroy refuses while pages are assigned
    with pytest.raises(_lib.HipError):
        a.destroy()
    a.deallocate(p2, 0)
    # page_allocator.rs:540-552: a reservation smaller than its pages
    with pytest.raises(_lib.HipError):
        a.allocate(ps, [2, 3], 2)
    # several addresses over distinct pages
    ptrs = [a.allocate(2 * ps, [2 * i + 2, 2 * i + 3], 10 + i) for i in range(3)]
    assert len(set(ptrs)) == 3
    for i, q in enumerate(ptrs):
        d = np.full(2 * ps, i + 1, np.uint8)
        assert (h2d_d2h(s, q, d) == d).all()
    for i, q in enumerate(ptrs):
        a.deallocate(q, 10 + i)
    a.destroy()
    s.destroy()


def test_page_allocator_unaligned_page_size():
    # page_allocator.rs:488-519
    g = H.page_granularity(0)
    if g == 1:
        pytest.skip("granularity 1: every page size is aligned")
    with pytest.raises(_lib.HipError):
        H.PageAllocator(0, g + 1, 4, True)


def test_device_facts():
    props = torch.cuda.get_device_properties(0)
    assert H.get_num_cus(0) == props.multi_processor_count
    assert H.arch_name(0) == props.gcnArchName.split(":")[0]


def test_file_to_device_and_back(tmp_path):
    s = H.HipStream(0)
    rng = np.random.default_rng(7)
    n = 5_000_003
    blob = rng.integers(0, 256, n + 4096, dtype=np.uint8)
    f = tmp_path / "data.bin"
    f.write_bytes(blob.tobytes())
    staging_bytes = 1 << 20
    st = H.alloc_pinned(staging_bytes)
    d = s.allocate(n)
    fd = H.open_file(str(f))
    H.read_file_to_device(fd, 4096, n, d, st, staging_bytes, s)
    back = np.zeros(n, np.uint8)
    s.memcpy_d2h(back.ctypes.data, d, n)
    s.sync()
    assert (back == blob[4096:]).all()
    import os
    os.close(fd)
    out = tmp_path / "out.bin"
    fd = H.open_file(str(out), write=True)
    H.write_file_from_device(fd, 100, n, d, st, staging_bytes, s)
    os.close(fd)
    raw = np.frombuffer(out.read_bytes(), np.uint8)
    assert raw.size == n + 100 and (raw[100:] == blob[4096:]).all()
    # reading past the end of the file is an error
    fd = H.open_file(str(f))
    with pytest.raises(_lib.HipError):
        H.read_file_to_device(fd, n, 8192, d, st, staging_bytes, s)
    os.close(fd)
    s.free(d)
    s.sync()
    H.free_pinned(st)
    s.destroy()
