// common.h -- shared host/device helpers for the HIP library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/impeller_hip.h"
#include "field.h"

namespace imp {

#define IMP_CHECK(expr)                                                                  \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) {                                                              \
      fprintf(stderr, "[impeller_hip] %s:%d: %s -> %s\n", __FILE__, __LINE__, #expr,     \
              hipGetErrorString(_e));                                                    \
      return _e;                                                                         \
    }                                                                                    \
  } while (0)

#define IMP_REQUIRE(cond, msg)                                                           \
  do {                                                                                   \
    if (!(cond)) {                                                                       \
      fprintf(stderr, "[impeller_hip] invalid argument: %s (%s)\n", msg, #cond);         \
      return hipErrorInvalidValue;                                                       \
    }                                                                                    \
  } while (0)

// A PolyPtr-style view: logical index i -> physical (off + i) mod wl, with
// off < wl and i < wl so one conditional subtraction suffices.
struct View {
  u32* base;
  u64 off;
  u64 wl;
};

__host__ __device__ __forceinline__ u64 view_phys(const View& v, u64 i) {
  u64 p = v.off + i;
  return p >= v.wl ? p - v.wl : p;
}

// iter.cuh:309-317: rotate != 0 requires an unsliced view and means
// offset = len - rotate.
inline bool make_view(const PolyPtr& p, View* v) {
  u64 off = p.offset, wl = p.whole_len ? p.whole_len : p.len;
  if (p.rotate != 0) {
    if (p.len != wl || p.offset != 0) return false;
    off = p.len - p.rotate;
  }
  if (p.len > wl || (wl && off >= wl)) return false;
  v->base = p.ptr;
  v->off = off;
  v->wl = wl;
  return true;
}

// Global-memory access through a pointer whose address space the compiler
// cannot see (one read from a job descriptor in memory): a plain access would
// be a flat one, which also counts against the LDS wait counter and so gets
// ordered against the kernel's LDS operations.  Every buffer the kernels
// address through such pointers is global memory.  (The host pass and the
// CPU emulation build see plain accesses.)
#if defined(IMP_EMU) || !defined(__HIP_DEVICE_COMPILE__)
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
  *p = v;
}
#else
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}
#endif

// A PolyPtr read from device memory as a view (make_slice_iter: a non-zero
// rotate means offset = len - rotate on an unsliced array)
__device__ __forceinline__ View dev_view(const PolyPtr* p) {
  const PolyPtr v = gld(p);
  View r;
  r.base = v.ptr;
  const u64 wl = v.whole_len ? v.whole_len : v.len;
  r.wl = wl ? wl : 1;
  r.off = v.rotate != 0 ? v.len - v.rotate : v.offset;
  return r;
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// The launch's dynamic LDS as a u32 pointer (one definition per kernel).
#ifndef IMP_DYNAMIC_LDS
#define IMP_DYNAMIC_LDS(name)            \
  extern __shared__ uint4 name##_lds4[]; \
  u32* name = reinterpret_cast<u32*>(name##_lds4)
#endif

__host__ __device__ __forceinline__ u64 bit_reverse(u64 x, u32 bits) {
  u64 r = 0;
  for (u32 i = 0; i < bits; i++) {
    r = (r << 1) | (x & 1);
    x >>= 1;
  }
  return r;
}

__device__ __forceinline__ u32 brev_bits(u32 x, u32 bits) {
  return bits == 0 ? 0u : (__builtin_bitreverse32(x) >> (32 - bits));
}

}  // namespace imp
