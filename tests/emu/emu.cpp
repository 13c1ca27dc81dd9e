// CPU emulation of HIP kernel launches (see hip/hip_runtime.h).  One
// work-group runs at a time; its work-items are ucontext fibers that yield at
// __syncthreads, and the group's barrier releases once every work-item that
// has not returned is waiting at it.  Dynamic LDS is a fresh heap block per
// work-group (so AddressSanitizer sees overruns); static __shared__ arrays are
// function-local statics.  Stream capture records launches and copies into a
// replayable list.  Test infrastructure only.
#include <stdlib.h>
#include <ucontext.h>

#include <vector>

#include "hip/hip_runtime.h"

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define EMU_ASAN 1
#include <sanitizer/common_interface_defs.h>
#endif
#endif

namespace emu {

dim3 threadIdx_, blockIdx_, blockDim_, gridDim_;

namespace {
const size_t kStack = 512 * 1024;

struct Fiber {
  ucontext_t ctx;
  char* stack = nullptr;
  int state = 0;  // 0 runnable, 1 waiting at the barrier, 2 returned, 3 in a wave collective
  int wkind = 0;  // collective: 1 ballot, 2 shuffle, 3 broadcast inside a group of four lanes
  unsigned long long wval = 0, wres = 0;
  int warg = 0;
#ifdef EMU_ASAN
  void* fake = nullptr;
#endif
};

std::vector<Fiber> g_fibers;
ucontext_t g_sched;
int g_cur = -1;
const std::function<void()>* g_body = nullptr;
char* g_lds = nullptr;
hipError_t g_last = hipSuccess;
bool g_capturing = false;
std::vector<std::function<void()>>* g_capture = nullptr;
#ifdef EMU_ASAN
const void* g_sched_bottom = nullptr;
size_t g_sched_size = 0;
void* g_sched_fake = nullptr;
#endif

// switch from the current fiber back to the scheduler
void to_sched(bool dying) {
  Fiber& f = g_fibers[g_cur];
#ifdef EMU_ASAN
  __sanitizer_start_switch_fiber(dying ? nullptr : &f.fake, g_sched_bottom, g_sched_size);
#else
  (void)dying;
#endif
  swapcontext(&f.ctx, &g_sched);
#ifdef EMU_ASAN
  __sanitizer_finish_switch_fiber(f.fake, &g_sched_bottom, &g_sched_size);
#endif
}

void trampoline() {
#ifdef EMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &g_sched_bottom, &g_sched_size);
#endif
  (*g_body)();
  g_fibers[g_cur].state = 2;
  to_sched(true);
}

void run_group(unsigned nthreads) {
  if (g_fibers.size() < nthreads) g_fibers.resize(nthreads);
  for (unsigned t = 0; t < nthreads; t++) {
    Fiber& f = g_fibers[t];
    if (!f.stack) f.stack = (char*)malloc(kStack);
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack;
    f.ctx.uc_stack.ss_size = kStack;
    f.ctx.uc_link = nullptr;
    makecontext(&f.ctx, trampoline, 0);
    f.state = 0;
  }
  for (;;) {
    for (unsigned t = 0; t < nthreads; t++) {
      Fiber& f = g_fibers[t];
      if (f.state != 0) continue;
      g_cur = (int)t;
      unsigned tx = t % blockDim_.x, ty = (t / blockDim_.x) % blockDim_.y,
               tz = t / (blockDim_.x * blockDim_.y);
      threadIdx_ = dim3(tx, ty, tz);
#ifdef EMU_ASAN
      __sanitizer_start_switch_fiber(&g_sched_fake, f.stack, kStack);
#endif
      swapcontext(&g_sched, &f.ctx);
#ifdef EMU_ASAN
      __sanitizer_finish_switch_fiber(g_sched_fake, nullptr, nullptr);
#endif
    }
    // group-of-four broadcasts: complete once every lane of the group that
    // has not returned is in one (the other lanes of the wave may be anywhere)
    bool released = false;
    for (unsigned q0 = 0; q0 < nthreads; q0 += 4) {
      unsigned q1 = q0 + 4 < nthreads ? q0 + 4 : nthreads;
      unsigned in_q = 0, live = 0;
      for (unsigned t = q0; t < q1; t++) {
        in_q += g_fibers[t].state == 3 && g_fibers[t].wkind == 3;
        live += g_fibers[t].state != 2;
      }
      if (in_q == 0) continue;
      if (in_q != live) {
        fprintf(stderr, "[emu] divergent group-of-four broadcast in block %u, lanes %u..%u:",
                blockIdx_.x, q0, q1 - 1);
        for (unsigned t = q0; t < q1; t++)
          fprintf(stderr, " %u:%d/%d", t, g_fibers[t].state, g_fibers[t].wkind);
        fprintf(stderr, "\n");
        abort();
      }
      for (unsigned t = q0; t < q1; t++) {
        Fiber& f = g_fibers[t];
        if (f.state != 3) continue;
        const Fiber& src = g_fibers[q0 + ((unsigned)f.warg & 3)];
        if (src.state != 3 || src.wkind != 3) {
          fprintf(stderr, "[emu] broadcast from a lane that has returned (block %u, lane %u)\n",
                  blockIdx_.x, q0 + ((unsigned)f.warg & 3));
          abort();
        }
        f.wres = src.wval;
      }
      for (unsigned t = q0; t < q1; t++)
        if (g_fibers[t].state == 3) g_fibers[t].state = 0;
      released = true;
    }
    if (released) continue;
    // wave collectives: complete once every lane of the wave that has not
    // returned is in one
    for (unsigned w0 = 0; w0 < nthreads; w0 += 64) {
      unsigned w1 = w0 + 64 < nthreads ? w0 + 64 : nthreads;
      unsigned in_op = 0, done = 0, at_bar = 0;
      for (unsigned t = w0; t < w1; t++) {
        in_op += g_fibers[t].state == 3;
        done += g_fibers[t].state == 2;
        at_bar += g_fibers[t].state == 1;
      }
      if (in_op == 0) continue;
      if (at_bar || in_op + done != w1 - w0) {
        fprintf(stderr, "[emu] divergent wave collective in block %u, wave %u\n", blockIdx_.x,
                w0 / 64);
        abort();
      }
      unsigned long long mask = 0;
      for (unsigned t = w0; t < w1; t++)
        if (g_fibers[t].state == 3 && g_fibers[t].wkind == 1 && g_fibers[t].wval)
          mask |= 1ull << (t - w0);
      for (unsigned t = w0; t < w1; t++) {
        Fiber& f = g_fibers[t];
        if (f.state != 3) continue;
        if (f.wkind == 1) f.wres = mask;
        else f.wres = g_fibers[w0 + ((unsigned)f.warg & 63)].wval;
        f.state = 0;
      }
      released = true;
    }
    if (released) continue;
    unsigned waiting = 0;
    for (unsigned t = 0; t < nthreads; t++) waiting += g_fibers[t].state == 1;
    if (waiting == 0) break;
    for (unsigned t = 0; t < nthreads; t++)
      if (g_fibers[t].state == 1) g_fibers[t].state = 0;
  }
  g_cur = -1;
}

void run_launch(dim3 grid, dim3 block, size_t shmem, const std::function<void()>& body) {
  unsigned nthreads = block.x * block.y * block.z;
  if (nthreads == 0 || nthreads > 1024 || grid.x == 0 || grid.y == 0 || grid.z == 0 ||
      shmem > 160 * 1024) {
    fprintf(stderr, "[emu] invalid launch: grid %u %u %u block %u %u %u lds %zu\n", grid.x,
            grid.y, grid.z, block.x, block.y, block.z, shmem);
    g_last = hipErrorInvalidValue;
    return;
  }
  const std::function<void()>* prev = g_body;
  g_body = &body;
  gridDim_ = grid;
  blockDim_ = block;
  for (unsigned z = 0; z < grid.z; z++)
    for (unsigned y = 0; y < grid.y; y++)
      for (unsigned x = 0; x < grid.x; x++) {
        blockIdx_ = dim3(x, y, z);
        g_lds = shmem ? (char*)malloc(shmem) : nullptr;
        if (g_lds) memset(g_lds, 0xA5, shmem);  // LDS starts undefined
        run_group(nthreads);
        free(g_lds);
        g_lds = nullptr;
      }
  g_body = prev;
}
}  // namespace

void syncthreads() {
  g_fibers[g_cur].state = 1;
  to_sched(false);
}

void* dyn_lds() { return g_lds; }

static unsigned long long wave_op(int kind, unsigned long long v, int arg) {
  Fiber& f = g_fibers[g_cur];
  f.wkind = kind;
  f.wval = v;
  f.warg = arg;
  f.state = 3;
  to_sched(false);
  return g_fibers[g_cur].wres;
}

unsigned long long wave_ballot(int pred) { return wave_op(1, pred ? 1 : 0, 0); }
unsigned long long wave_shfl(unsigned long long v, int src) { return wave_op(2, v, src); }
unsigned long long quad_shfl(unsigned long long v, int k) { return wave_op(3, v, k); }
unsigned lane_id() { return (unsigned)g_cur & 63u; }

void launch_impl(dim3 grid, dim3 block, size_t shmem, std::function<void()> body) {
  if (g_capturing) {
    g_capture->push_back([grid, block, shmem, body]() { run_launch(grid, block, shmem, body); });
    return;
  }
  run_launch(grid, block, shmem, body);
}

bool capturing() { return g_capturing; }
void record(std::function<void()> f) { g_capture->push_back(std::move(f)); }

}  // namespace emu

using namespace emu;

struct ihipGraph_t {
  std::vector<std::function<void()>> ops;
};
struct ihipGraphExec_t {
  std::vector<std::function<void()>> ops;
};
struct ihipEvent_t {
  int dummy;
};

hipError_t hipMalloc(void** p, size_t n) {
  *p = malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  free(p);
  return hipSuccess;
}
hipError_t hipMallocAsync(void** p, size_t n, hipStream_t) {
  if (g_capturing) return hipErrorNotSupported;
  return hipMalloc(p, n);
}
hipError_t hipFreeAsync(void* p, hipStream_t) {
  if (g_capturing) return hipErrorNotSupported;
  return hipFree(p);
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  if (g_capturing) return hipErrorNotSupported;
  memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  if (g_capturing) {
    g_capture->push_back([d, s, n]() { memmove(d, s, n); });
    return hipSuccess;
  }
  memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
  // a graph must never hold a memset node (clear with a kernel instead)
  if (g_capturing) return hipErrorNotSupported;
  memset(d, v, n);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) {
  return g_capturing ? hipErrorNotSupported : hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() {
  hipError_t e = g_last;
  g_last = hipSuccess;
  return e;
}
const char* hipGetErrorString(hipError_t e) {
  switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorInvalidValue: return "hipErrorInvalidValue";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
    case hipErrorNotSupported: return "hipErrorNotSupported";
    default: return "hipError (emulated)";
  }
}
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
// the emulated allocator has no pool: the attribute is accepted and ignored
hipError_t hipDeviceGetDefaultMemPool(hipMemPool_t* pool, int) {
  *pool = nullptr;
  return hipSuccess;
}
hipError_t hipMemPoolSetAttribute(hipMemPool_t, hipMemPoolAttr, void*) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
  *v = 8;  // a small "chip": persistent grids stride more often
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  static int token;
  *s = (hipStream_t)&token;
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new ihipEvent_t{0};
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
  *p = const_cast<void*>(symbol);
  return hipSuccess;
}
// launches run to completion in issue order, so every wait is already met
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.f;
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  if (g_capturing) return hipErrorInvalidValue;
  g_capturing = true;
  g_capture = new std::vector<std::function<void()>>();
  return hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* status) {
  *status = g_capturing ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  if (!g_capturing) return hipErrorInvalidValue;
  g_capturing = false;
  *g = new ihipGraph_t{std::move(*g_capture)};
  delete g_capture;
  g_capture = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *e = new ihipGraphExec_t{g->ops};
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& op : e->ops) op();
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
