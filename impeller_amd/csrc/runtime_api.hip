// runtime_api.hip -- the device-runtime layer the reference's Rust runtime
// links against (its zkpoly_cuda_api crate: cuda_api/src/stream.rs, mem.rs,
// mem/static_allocator.rs, mem/page_allocator.rs, device_info.rs, file.rs),
// restated over the HIP runtime behind a C ABI (impeller_rt_*), so a
// hip_api crate binds these symbols instead of generating CUDA bindings.
// Host code only: streams, host-ordered events, pinned host memory, a
// caller-planned static allocator, a page allocator over HIP virtual memory
// (hipMemCreate / hipMemMap), device facts, and file <-> device transfers
// through POSIX reads and writes (the reference uses cuFile / GPUDirect
// Storage, which has no ROCm counterpart here).
//
// Errors are returned, never raised: a hipError_t from the runtime, or
// hipErrorInvalidValue (with a message on stderr) for a broken contract the
// reference answers with a panic (overlapping static allocations, a page
// assigned twice, pages still mapped at destruction).
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <iterator>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.h"

using imp::u32;
using imp::u64;

#define RT_CHECK(expr)                                                                   \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) {                                                              \
      fprintf(stderr, "[impeller_rt] %s:%d: %s -> %s\n", __FILE__, __LINE__, #expr,      \
              hipGetErrorString(_e));                                                    \
      return _e;                                                                         \
    }                                                                                    \
  } while (0)

#define RT_REQUIRE(cond, ...)                                                            \
  do {                                                                                   \
    if (!(cond)) {                                                                       \
      fprintf(stderr, "[impeller_rt] ");                                                 \
      fprintf(stderr, __VA_ARGS__);                                                      \
      fprintf(stderr, "\n");                                                             \
      return hipErrorInvalidValue;                                                       \
    }                                                                                    \
  } while (0)

// ------------------------------------------------------------------ streams
// cuda_api/src/stream.rs:137-257 (CudaStream)

extern "C" hipError_t impeller_rt_stream_create(int device, hipStream_t* out) {
  RT_REQUIRE(out, "null stream out-pointer");
  RT_CHECK(hipSetDevice(device));
  return hipStreamCreate(out);
}

extern "C" hipError_t impeller_rt_stream_destroy(hipStream_t s) { return hipStreamDestroy(s); }

extern "C" hipError_t impeller_rt_stream_sync(hipStream_t s) { return hipStreamSynchronize(s); }

// stream-ordered device memory (cudaMallocAsync / cudaFreeAsync)
extern "C" hipError_t impeller_rt_stream_alloc(hipStream_t s, size_t bytes, void** out) {
  RT_REQUIRE(out, "null out-pointer");
  return hipMallocAsync(out, bytes, s);
}

extern "C" hipError_t impeller_rt_stream_free(hipStream_t s, void* p) { return hipFreeAsync(p, s); }

// kind: hipMemcpyHostToDevice (1), DeviceToHost (2), DeviceToDevice (3)
extern "C" hipError_t impeller_rt_memcpy_async(int device, void* dst, const void* src, size_t bytes,
                                               int kind, hipStream_t s) {
  RT_REQUIRE(kind >= 1 && kind <= 3, "memcpy kind %d (1 h2d, 2 d2h, 3 d2d)", kind);
  RT_CHECK(hipSetDevice(device));
  return hipMemcpyAsync(dst, src, bytes, (hipMemcpyKind)kind, s);
}

// ------------------------------------------------------------------- events
// cuda_api/src/stream.rs:1-135.  A host-ordered event (CudaEvent) can be
// waited on from another host thread before it is recorded: sync and
// stream_wait first block until some thread has recorded it (a flag under a
// mutex and condition variable), so the wait always refers to the record
// meant by the program, not to an earlier state of the event.  A raw event
// (CudaEventRaw) is the plain runtime event.  Both use blocking sync.
struct ImpRtEvent {
  hipEvent_t ev = nullptr;
  int device = 0;
  bool host_ordered = false;
  std::mutex m;
  std::condition_variable cv;
  bool recorded = false;
};

extern "C" hipError_t impeller_rt_event_create(int device, int host_ordered, ImpRtEvent** out) {
  RT_REQUIRE(out, "null event out-pointer");
  RT_CHECK(hipSetDevice(device));
  ImpRtEvent* e = new ImpRtEvent;
  e->device = device;
  e->host_ordered = host_ordered != 0;
  hipError_t r = hipEventCreateWithFlags(&e->ev, hipEventBlockingSync);
  if (r != hipSuccess) {
    delete e;
    return r;
  }
  *out = e;
  return hipSuccess;
}

static void wait_recorded(ImpRtEvent* e) {
  if (!e->host_ordered) return;
  std::unique_lock<std::mutex> lk(e->m);
  e->cv.wait(lk, [e] { return e->recorded; });
}

extern "C" hipError_t impeller_rt_event_record(ImpRtEvent* e, hipStream_t s) {
  RT_REQUIRE(e, "null event");
  RT_CHECK(hipSetDevice(e->device));
  RT_CHECK(hipEventRecord(e->ev, s));
  if (e->host_ordered) {
    std::lock_guard<std::mutex> lk(e->m);
    e->recorded = true;
    e->cv.notify_all();
  }
  return hipSuccess;
}

extern "C" hipError_t impeller_rt_event_sync(ImpRtEvent* e) {
  RT_REQUIRE(e, "null event");
  wait_recorded(e);
  RT_CHECK(hipSetDevice(e->device));
  return hipEventSynchronize(e->ev);
}

// the stream waits on the event (stream.rs CudaStream::wait / wait_raw)
extern "C" hipError_t impeller_rt_stream_wait_event(hipStream_t s, ImpRtEvent* e) {
  RT_REQUIRE(e, "null event");
  wait_recorded(e);
  return hipStreamWaitEvent(s, e->ev, 0);
}

// forget the record, so the next sync waits for a new one (CudaEvent::reset)
extern "C" hipError_t impeller_rt_event_reset(ImpRtEvent* e) {
  RT_REQUIRE(e, "null event");
  std::lock_guard<std::mutex> lk(e->m);
  e->recorded = false;
  return hipSuccess;
}

// milliseconds from a to b (both recorded and complete)
extern "C" hipError_t impeller_rt_event_elapsed(ImpRtEvent* a, ImpRtEvent* b, float* ms) {
  RT_REQUIRE(a && b && ms, "null argument");
  RT_CHECK(hipSetDevice(a->device));
  return hipEventElapsedTime(ms, a->ev, b->ev);
}

extern "C" hipEvent_t impeller_rt_event_raw(ImpRtEvent* e) { return e ? e->ev : nullptr; }

extern "C" hipError_t impeller_rt_event_destroy(ImpRtEvent* e) {
  if (!e) return hipSuccess;
  hipError_t r = hipSetDevice(e->device);
  if (r == hipSuccess) r = hipEventDestroy(e->ev);
  delete e;
  return r;
}

// -------------------------------------------------------- pinned host memory
// cuda_api/src/mem.rs:18-46 (alloc_pinned / free_pinned)

extern "C" hipError_t impeller_rt_alloc_pinned(size_t bytes, void** out) {
  RT_REQUIRE(out, "null out-pointer");
  return hipHostMalloc(out, bytes, hipHostMallocDefault);
}

extern "C" hipError_t impeller_rt_free_pinned(void* p) { return hipHostFree(p); }

// --------------------------------------------------------- static allocator
// cuda_api/src/mem/static_allocator.rs: one device allocation of max_size
// bytes; the caller (the reference's memory planner) chooses every offset.
// With check_overlap the allocator keeps the live ranges [left, right) and
// refuses a range that meets one of them.
struct ImpRtStatic {
  int device = 0;
  char* base = nullptr;
  size_t max_size = 0;
  bool check = false;
  std::map<size_t, size_t> ranges;  // left -> right
};

extern "C" hipError_t impeller_rt_static_create(int device, size_t max_size, int check_overlap,
                                                ImpRtStatic** out) {
  RT_REQUIRE(out && max_size > 0, "null out-pointer or zero size");
  RT_CHECK(hipSetDevice(device));
  void* p = nullptr;
  RT_CHECK(hipMalloc(&p, max_size));
  ImpRtStatic* a = new ImpRtStatic;
  a->device = device;
  a->base = static_cast<char*>(p);
  a->max_size = max_size;
  a->check = check_overlap != 0;
  *out = a;
  return hipSuccess;
}

extern "C" hipError_t impeller_rt_static_allocate(ImpRtStatic* a, size_t offset, size_t bytes,
                                                  void** out) {
  RT_REQUIRE(a && out, "null argument");
  RT_REQUIRE(offset < a->max_size && bytes <= a->max_size - offset,
             "range [%zu, %zu) outside the %zu-byte arena", offset, offset + bytes, a->max_size);
  const size_t left = offset, right = offset + bytes;
  if (a->check) {
    auto next = a->ranges.lower_bound(left);
    if (next != a->ranges.begin()) {
      auto prev = std::prev(next);
      RT_REQUIRE(prev->second <= left, "overlap: [%zu, %zu) meets [%zu, %zu)", left, right,
                 prev->first, prev->second);
    }
    if (next != a->ranges.end())
      RT_REQUIRE(next->first >= right, "overlap: [%zu, %zu) meets [%zu, %zu)", left, right,
                 next->first, next->second);
    a->ranges[left] = right;
  }
  *out = a->base + offset;
  return hipSuccess;
}

extern "C" hipError_t impeller_rt_static_free(ImpRtStatic* a, void* p) {
  RT_REQUIRE(a, "null allocator");
  if (!a->check) return hipSuccess;
  RT_REQUIRE(p >= (void*)a->base && p < (void*)(a->base + a->max_size), "pointer outside the arena");
  const size_t off = static_cast<size_t>(static_cast<char*>(p) - a->base);
  auto it = a->ranges.find(off);
  RT_REQUIRE(it != a->ranges.end(), "free of offset %zu, which holds no allocation", off);
  a->ranges.erase(it);
  return hipSuccess;
}

extern "C" void* impeller_rt_static_base(ImpRtStatic* a) { return a ? a->base : nullptr; }

extern "C" hipError_t impeller_rt_static_destroy(ImpRtStatic* a) {
  if (!a) return hipSuccess;
  hipError_t r = hipSetDevice(a->device);
  if (r == hipSuccess) r = hipFree(a->base);
  delete a;
  return r;
}

// ----------------------------------------------------------- page allocator
// cuda_api/src/mem/page_allocator.rs over HIP virtual memory: page_num
// physical pages of page_size bytes (a multiple of the minimum granularity)
// are created up front; allocate(va_size, page_ids, vaid) reserves va_size
// bytes of address space the first time a vaid is seen and maps the listed
// pages there in order; later calls with the same vaid return the same
// address (the mapping stays until the allocator is destroyed, as in the
// reference).  With memory_check, a page may belong to one live address at a
// time, a later allocation of a vaid must name the pages mapped there,
// deallocate releases the pages of an address, and destroy refuses to run
// while pages are still assigned.
struct ImpRtVa {
  char* ptr = nullptr;
  size_t size = 0;
  std::vector<size_t> mapped;    // the pages mapped there, in order
  std::vector<size_t> assigned;  // pages currently assigned (memory_check)
};

struct ImpRtPages {
  int device = 0;
  size_t page_size = 0;
  std::vector<hipMemGenericAllocationHandle_t> table;
  bool check = false;
  std::vector<char*> owner;  // page -> the address it is assigned to (memory_check)
  std::unordered_map<size_t, ImpRtVa> vas;
};

static hipMemAllocationProp device_prop(int device) {
  hipMemAllocationProp prop = {};
  prop.type = hipMemAllocationTypePinned;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = device;
  return prop;
}

extern "C" hipError_t impeller_rt_page_granularity(int device, size_t* out) {
  RT_REQUIRE(out, "null out-pointer");
  RT_CHECK(hipSetDevice(device));
  int vmm = 0;
  RT_CHECK(hipDeviceGetAttribute(&vmm, hipDeviceAttributeVirtualMemoryManagementSupported, device));
  if (!vmm) return hipErrorNotSupported;
  hipMemAllocationProp prop = device_prop(device);
  return hipMemGetAllocationGranularity(out, &prop, hipMemAllocationGranularityMinimum);
}

extern "C" hipError_t impeller_rt_page_destroy(ImpRtPages* a);

extern "C" hipError_t impeller_rt_page_create(int device, size_t page_size, size_t page_num,
                                              int memory_check, ImpRtPages** out) {
  RT_REQUIRE(out && page_size > 0, "null out-pointer or zero page size");
  size_t gran = 0;
  RT_CHECK(impeller_rt_page_granularity(device, &gran));
  RT_REQUIRE(gran > 0 && page_size % gran == 0,
             "page size must be aligned to minimum page size %zu, but got %zu", gran, page_size);
  ImpRtPages* a = new ImpRtPages;
  a->device = device;
  a->page_size = page_size;
  a->check = memory_check != 0;
  a->owner.assign(a->check ? page_num : 0, nullptr);
  hipMemAllocationProp prop = device_prop(device);
  for (size_t i = 0; i < page_num; i++) {
    hipMemGenericAllocationHandle_t h;
    hipError_t r = hipMemCreate(&h, page_size, &prop, 0);
    if (r != hipSuccess) {
      fprintf(stderr, "[impeller_rt] hipMemCreate of page %zu -> %s\n", i, hipGetErrorString(r));
      impeller_rt_page_destroy(a);
      return r;
    }
    a->table.push_back(h);
  }
  *out = a;
  return hipSuccess;
}

extern "C" hipError_t impeller_rt_page_allocate(ImpRtPages* a, size_t va_size, const size_t* page_ids,
                                                size_t n_pages, size_t vaid, void** out) {
  RT_REQUIRE(a && out && (page_ids || n_pages == 0), "null argument");
  for (size_t k = 0; k < n_pages; k++)
    RT_REQUIRE(page_ids[k] < a->table.size(), "page id %zu of %zu pages", page_ids[k],
               a->table.size());
  auto found = a->vas.find(vaid);
  if (found != a->vas.end()) {
    ImpRtVa& va = found->second;
    if (a->check) {
      RT_REQUIRE(va_size == va.size, "inconsistent va size %zu for vaid %zu (reserved %zu)",
                 va_size, vaid, va.size);
      // the address keeps the pages of its first allocation: a later one
      // must name those pages (another list would alias other memory)
      RT_REQUIRE(n_pages == va.mapped.size() &&
                     std::equal(page_ids, page_ids + n_pages, va.mapped.begin()),
                 "vaid %zu is mapped to other pages", vaid);
      for (size_t k = 0; k < n_pages; k++)
        RT_REQUIRE(a->owner[page_ids[k]] == nullptr,
                   "allocating page %zu to virtual address %p while it is still occupied by %p",
                   page_ids[k], (void*)va.ptr, (void*)a->owner[page_ids[k]]);
      for (size_t k = 0; k < n_pages; k++) {
        a->owner[page_ids[k]] = va.ptr;
        va.assigned.push_back(page_ids[k]);
      }
    }
    *out = va.ptr;
    return hipSuccess;
  }
  RT_REQUIRE(va_size >= a->page_size * n_pages, "va_size %zu < %zu pages of %zu bytes", va_size,
             n_pages, a->page_size);
  if (a->check)
    for (size_t k = 0; k < n_pages; k++)
      RT_REQUIRE(a->owner[page_ids[k]] == nullptr,
                 "allocating page %zu while it is still occupied by %p", page_ids[k],
                 (void*)a->owner[page_ids[k]]);
  RT_CHECK(hipSetDevice(a->device));
  void* base = nullptr;
  RT_CHECK(hipMemAddressReserve(&base, va_size, a->page_size, nullptr, 0));
  char* p = static_cast<char*>(base);
  size_t mapped = 0;
  hipError_t r = hipSuccess;
  while (mapped < n_pages) {
    r = hipMemMap(p + mapped * a->page_size, a->page_size, 0, a->table[page_ids[mapped]], 0);
    if (r != hipSuccess) break;
    mapped++;
  }
  if (r == hipSuccess && n_pages) {
    hipMemAccessDesc acc = {};
    acc.location.type = hipMemLocationTypeDevice;
    acc.location.id = a->device;
    acc.flags = hipMemAccessFlagsProtReadWrite;
    r = hipMemSetAccess(p, n_pages * a->page_size, &acc, 1);
  }
  if (r != hipSuccess) {
    fprintf(stderr, "[impeller_rt] mapping vaid %zu -> %s\n", vaid, hipGetErrorString(r));
    for (size_t k = 0; k < mapped; k++) hipMemUnmap(p + k * a->page_size, a->page_size);
    hipMemAddressFree(base, va_size);
    return r;
  }
  ImpRtVa va;
  va.ptr = p;
  va.size = va_size;
  va.mapped.assign(page_ids, page_ids + n_pages);
  if (a->check)
    for (size_t k = 0; k < n_pages; k++) {
      a->owner[page_ids[k]] = p;
      va.assigned.push_back(page_ids[k]);
    }
  a->vas[vaid] = va;
  *out = p;
  return hipSuccess;
}

extern "C" hipError_t impeller_rt_page_deallocate(ImpRtPages* a, void* p, size_t vaid) {
  RT_REQUIRE(a, "null allocator");
  if (!a->check) return hipSuccess;
  auto found = a->vas.find(vaid);
  RT_REQUIRE(found != a->vas.end(), "cannot deallocate unallocated vaid %zu", vaid);
  ImpRtVa& va = found->second;
  RT_REQUIRE(va.ptr == p, "vaid %zu inconsistent with pointer %p", vaid, p);
  RT_REQUIRE(!va.assigned.empty(), "virtual address %p is not mapped to any physical page", p);
  for (size_t id : va.assigned) {
    RT_REQUIRE(a->owner[id] == va.ptr, "freeing page %zu for %p, but it is mapped to %p", id, p,
               (void*)a->owner[id]);
    a->owner[id] = nullptr;
  }
  va.assigned.clear();
  return hipSuccess;
}

extern "C" size_t impeller_rt_page_size(ImpRtPages* a) { return a ? a->page_size : 0; }

extern "C" hipError_t impeller_rt_page_destroy(ImpRtPages* a) {
  if (!a) return hipSuccess;
  if (a->check)
    for (size_t i = 0; i < a->owner.size(); i++)
      RT_REQUIRE(a->owner[i] == nullptr,
                 "upon destroying the page allocator, page %zu is still mapped to %p", i,
                 (void*)a->owner[i]);
  hipError_t first = hipSetDevice(a->device);
  auto keep = [&](hipError_t r) {
    if (first == hipSuccess) first = r;
  };
  for (auto& kv : a->vas) {
    for (size_t k = 0; k < kv.second.mapped.size(); k++)
      keep(hipMemUnmap(kv.second.ptr + k * a->page_size, a->page_size));
    keep(hipMemAddressFree(kv.second.ptr, kv.second.size));
  }
  for (auto h : a->table) keep(hipMemRelease(h));
  delete a;
  return first;
}

// -------------------------------------------------------------- device info
// cuda_api/src/device_info.rs: the compute-unit count (get_num_sms) and the
// architecture name (sm_version; here the gfx target, e.g. "gfx950")

extern "C" hipError_t impeller_rt_num_cus(int device, int* out) {
  RT_REQUIRE(out, "null out-pointer");
  return hipDeviceGetAttribute(out, hipDeviceAttributeMultiprocessorCount, device);
}

extern "C" hipError_t impeller_rt_arch_name(int device, char* buf, size_t len) {
  RT_REQUIRE(buf && len > 0, "null buffer");
  hipDeviceProp_t prop;
  RT_CHECK(hipGetDeviceProperties(&prop, device));
  const char* name = prop.gcnArchName;
  size_t k = 0;
  for (; k + 1 < len && name[k] && name[k] != ':'; k++) buf[k] = name[k];
  buf[k] = 0;
  return hipSuccess;
}

// ---------------------------------------------------------------- file I/O
// cuda_api/src/file.rs registers a file with cuFile for direct reads into
// device memory.  Here the bytes [file_offset, file_offset + bytes) move
// between a file descriptor and device memory through a caller-provided
// pinned staging buffer, split in two halves so the read of one chunk
// overlaps the copy of the previous one.  Both calls return once the data is
// in place (device memory for a read, the file for a write), as cuFileRead /
// cuFileWrite do.

extern "C" hipError_t impeller_rt_file_read_to_device(int fd, u64 file_offset, size_t bytes,
                                                      void* dst, void* staging,
                                                      size_t staging_bytes, hipStream_t s) {
  RT_REQUIRE(fd >= 0 && (dst || bytes == 0) && staging && staging_bytes >= 2,
             "bad file-read argument");
  const size_t half = staging_bytes / 2;
  char* stage[2] = {static_cast<char*>(staging), static_cast<char*>(staging) + half};
  hipEvent_t done[2] = {nullptr, nullptr};
  hipError_t r = hipSuccess;
  for (int k = 0; k < 2 && r == hipSuccess; k++) r = hipEventCreateWithFlags(&done[k], hipEventDisableTiming);
  size_t pos = 0;
  for (u32 i = 0; r == hipSuccess && pos < bytes; i++) {
    const u32 b = i & 1;
    const size_t n = bytes - pos < half ? bytes - pos : half;
    if (i >= 2) r = hipEventSynchronize(done[b]);  // the copy that last used this half
    if (r != hipSuccess) break;
    size_t got = 0;
    while (got < n) {
      const ssize_t k = pread(fd, stage[b] + got, n - got, (off_t)(file_offset + pos + got));
      if (k <= 0) {
        fprintf(stderr, "[impeller_rt] pread at %llu: %s\n",
                (unsigned long long)(file_offset + pos + got), k == 0 ? "end of file" : "error");
        r = hipErrorInvalidValue;
        break;
      }
      got += (size_t)k;
    }
    if (r != hipSuccess) break;
    r = hipMemcpyAsync(static_cast<char*>(dst) + pos, stage[b], n, hipMemcpyHostToDevice, s);
    if (r == hipSuccess) r = hipEventRecord(done[b], s);
    pos += n;
  }
  hipError_t w = hipStreamSynchronize(s);
  for (int k = 0; k < 2; k++)
    if (done[k]) hipEventDestroy(done[k]);
  return r != hipSuccess ? r : w;
}

extern "C" hipError_t impeller_rt_file_write_from_device(int fd, u64 file_offset, size_t bytes,
                                                         const void* src, void* staging,
                                                         size_t staging_bytes, hipStream_t s) {
  RT_REQUIRE(fd >= 0 && (src || bytes == 0) && staging && staging_bytes >= 1,
             "bad file-write argument");
  char* stage = static_cast<char*>(staging);
  for (size_t pos = 0; pos < bytes;) {
    const size_t n = bytes - pos < staging_bytes ? bytes - pos : staging_bytes;
    RT_CHECK(hipMemcpyAsync(stage, static_cast<const char*>(src) + pos, n, hipMemcpyDeviceToHost, s));
    RT_CHECK(hipStreamSynchronize(s));
    size_t put = 0;
    while (put < n) {
      const ssize_t k = pwrite(fd, stage + put, n - put, (off_t)(file_offset + pos + put));
      RT_REQUIRE(k > 0, "pwrite at %llu failed", (unsigned long long)(file_offset + pos + put));
      put += (size_t)k;
    }
    pos += n;
  }
  return hipSuccess;
}
