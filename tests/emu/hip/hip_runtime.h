// CPU emulation of the small part of the HIP runtime and kernel language the
// library uses, so that its kernels can run on the host (optionally under
// AddressSanitizer) for memory-safety and logic checks.  Kernels run one
// work-group at a time; the work-items of a group are fibers that switch at
// __syncthreads.  "Device" memory is host memory.  Test infrastructure only.
#pragma once
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <functional>
#include <memory>
#include <tuple>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint4 {
  unsigned x, y, z, w;
};
inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) { return uint4{a, b, c, d}; }

enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorLaunchFailure = 719,
  hipErrorNotSupported = 801,
  hipErrorUnknown = 999
};
enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
};
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };
enum hipStreamCaptureStatus { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive = 1 };
#define hipStreamNonBlocking 0x1

typedef struct ihipStream_t* hipStream_t;
typedef struct ihipEvent_t* hipEvent_t;
typedef struct ihipGraph_t* hipGraph_t;
typedef struct ihipGraphExec_t* hipGraphExec_t;
typedef struct ihipGraphNode_t* hipGraphNode_t;

namespace emu {
extern dim3 threadIdx_, blockIdx_, blockDim_, gridDim_;
void syncthreads();
// wave-wide collectives: every lane of the wave that has not returned takes part
unsigned long long wave_ballot(int pred);
unsigned long long wave_shfl(unsigned long long v, int src);
// v of lane 4 * (lane / 4) + k, every lane of that group of four taking part
unsigned long long quad_shfl(unsigned long long v, int k);
unsigned lane_id();
void* dyn_lds();
// runs body for every work-item of the grid; while a stream capture is open
// the launch is recorded instead (the body owns copies of its arguments)
void launch_impl(dim3 grid, dim3 block, size_t shmem, std::function<void()> body);

template <class... P, class... A>
inline void launch(void (*k)(P...), dim3 g, dim3 b, size_t sh, hipStream_t, A&&... a) {
  auto args = std::make_shared<std::tuple<P...>>(static_cast<P>(a)...);
  launch_impl(g, b, sh, [k, args]() { std::apply(k, *args); });
}
}  // namespace emu

#define threadIdx (::emu::threadIdx_)
#define blockIdx (::emu::blockIdx_)
#define blockDim (::emu::blockDim_)
#define gridDim (::emu::gridDim_)
#define __syncthreads() ::emu::syncthreads()
#define IMP_DYNAMIC_LDS(name) u32* name = reinterpret_cast<u32*>(::emu::dyn_lds())
#define hipLaunchKernelGGL(K, G, B, S, ST, ...) \
  ::emu::launch(K, dim3(G), dim3(B), (size_t)(S), ST, ##__VA_ARGS__)

inline unsigned atomicAdd(unsigned* p, unsigned v) {
  return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
}
inline unsigned atomicAnd(unsigned* p, unsigned v) {
  return __atomic_fetch_and(p, v, __ATOMIC_RELAXED);
}
inline unsigned atomicOr(unsigned* p, unsigned v) {
  return __atomic_fetch_or(p, v, __ATOMIC_RELAXED);
}
inline unsigned atomicMax(unsigned* p, unsigned v) {
  unsigned cur = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
  return cur;
}
inline unsigned long long __ballot(int pred) { return ::emu::wave_ballot(pred); }
inline unsigned __lane_id() { return ::emu::lane_id(); }
inline int __shfl(int v, int src, int = 64) {
  return (int)(unsigned)::emu::wave_shfl((unsigned)v, src);
}
inline unsigned __shfl(unsigned v, int src, int = 64) {
  return (unsigned)::emu::wave_shfl(v, src);
}
inline int __shfl_xor(int v, int mask, int = 64) {
  return (int)(unsigned)::emu::wave_shfl((unsigned)v, (int)(::emu::lane_id() ^ (unsigned)mask));
}
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline unsigned __popcll(unsigned long long x) { return (unsigned)__builtin_popcountll(x); }

hipError_t hipMalloc(void** p, size_t n);
hipError_t hipFree(void* p);
hipError_t hipMallocAsync(void** p, size_t n, hipStream_t s);
hipError_t hipFreeAsync(void* p, hipStream_t s);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetDevice(int* d);
hipError_t hipSetDevice(int d);
typedef struct EmuMemPool* hipMemPool_t;
enum hipMemPoolAttr { hipMemPoolAttrReleaseThreshold = 4 };
hipError_t hipDeviceGetDefaultMemPool(hipMemPool_t* pool, int dev);
hipError_t hipMemPoolSetAttribute(hipMemPool_t pool, hipMemPoolAttr attr, void* value);
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int dev);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipEventCreate(hipEvent_t* e);
#define hipEventDisableTiming 0x2
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
// a __device__ variable is host memory here: its address is the symbol's
#define HIP_SYMBOL(x) (reinterpret_cast<const void*>(&(x)))
hipError_t hipGetSymbolAddress(void** p, const void* symbol);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus* status);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* n, char* log,
                               size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
