// msm.hip -- BN254 G1 multi-scalar multiplication (Pippenger) for gfx950.
//
// Semantics follow the reference MSM (core/src/msm/src/msm_impl.cuh and
// msm.cuh:51-70): scalars are Fr in Montgomery form, converted to integers,
// cut into W = floor(254 / c) + 1 signed c-bit digits (|digit| <= 2^(c-1), so
// 2^(c-1) buckets per window).  The reference's precompute option is kept:
// with n_windows physical windows, logical window w uses point table
// w / n_windows (P * 2^((w / n_windows) * n_windows * c)) and accumulates into
// physical window w % n_windows (msm_impl.cuh:56-67), so n_windows = 1 needs
// no doublings at the end.  The result is normalised to affine, which is
// unique, so it is bit-identical to any correct CPU MSM on the same inputs.
//
// Pipeline (all on one stream, device resident):
//   k_msm_digits   scalar -> signed digits, per-bucket histogram (atomics)
//   k_msm_scan     per-window exclusive scan of the histogram
//   k_msm_scatter  counting-sort scatter of (point index | sign) by bucket
//   k_msm_buckets  mixed XYZZ+affine accumulation, one thread per bucket run
//   k_msm_reduce   sum_b b * B_b via running sums over bucket segments
//   k_msm_final    window partials -> Horner over windows -> affine
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.h"
#include "timing.h"

namespace imp {

struct MsmShape {
  u32 c;           // window bits
  u32 W;           // logical windows
  u32 n_windows;   // physical windows
  u32 n_pre;       // point tables
  u32 nb;          // buckets per physical window = 2^(c-1)
  u32 seg;         // buckets per reduce segment
  u32 nseg;        // segments per physical window
  u64 N;           // points per table
};

static MsmShape msm_shape(u64 N, u32 c, u32 target_windows) {
  MsmShape s;
  s.c = c;
  s.W = 254 / c + 1;
  s.n_windows = target_windows == 0 || target_windows > s.W ? s.W : target_windows;
  s.n_pre = (s.W + s.n_windows - 1) / s.n_windows;
  s.nb = 1u << (c - 1);
  s.seg = s.nb >= 8 ? 8 : s.nb;
  s.nseg = s.nb / s.seg;
  s.N = N;
  return s;
}

struct MsmLayout {
  size_t keys, sorted, counts, offsets, cursor, buckets, partials, wsum, total;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static MsmLayout msm_layout(const MsmShape& s) {
  MsmLayout l;
  size_t o = 0;
  l.keys = o;
  o = align256(o + (size_t)s.W * s.N * 4);
  l.sorted = o;
  o = align256(o + (size_t)s.n_windows * s.n_pre * s.N * 4);
  l.counts = o;
  o = align256(o + (size_t)s.n_windows * (s.nb + 1) * 4);
  l.offsets = o;
  o = align256(o + (size_t)s.n_windows * (s.nb + 2) * 4);
  l.cursor = o;
  o = align256(o + (size_t)s.n_windows * (s.nb + 1) * 4);
  l.buckets = o;
  o = align256(o + (size_t)s.n_windows * s.nb * 128);
  l.partials = o;
  o = align256(o + (size_t)s.n_windows * s.nseg * 128);
  l.wsum = o;
  o = align256(o + (size_t)s.n_windows * 128);
  l.total = o;
  return l;
}

// ------------------------------------------------------------------ kernels
__global__ void k_zero_u32(u32* p, u64 n) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// signed digits (msm_impl.cuh:22-33 with digit == 2^(c-1) kept positive, so
// the top window never carries out): key = |digit| | sign << 31
__global__ void k_msm_digits(const u32* scalars, MsmShape s, u32* keys, u32* counts) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.N) return;
  Fr v = to_number<FrParams>(load_fe<Fr>(scalars + 8 * i));
  const u32 c = s.c, half = 1u << (c - 1), mask = (1u << c) - 1;
  u32 carry = 0;
  for (u32 w = 0; w < s.W; w++) {
    u32 bit = w * c, limb = bit >> 5, off = bit & 31;
    u64 word = v.v[limb];
    if (limb + 1 < 8) word |= (u64)v.v[limb + 1] << 32;
    u32 d = (u32)((word >> off) & mask) + carry;
    u32 key;
    if (d > half) {
      key = ((1u << c) - d) | 0x80000000u;
      carry = 1;
    } else {
      key = d;
      carry = 0;
    }
    keys[(u64)w * s.N + i] = key;
    u32 b = key & 0x7fffffffu;
    if (b) atomicAdd(&counts[(w % s.n_windows) * (s.nb + 1) + b], 1u);
  }
}

// one block per physical window: offsets[b] = sum_{b' < b} counts[b'],
// offsets[nb + 1] = total; cursor = offsets
__global__ __launch_bounds__(1024) void k_msm_scan(MsmShape s, const u32* counts, u32* offsets,
                                                   u32* cursor) {
  __shared__ u32 part[1024];
  const u32 pw = blockIdx.x, n = s.nb + 1;
  const u32* cnt = counts + pw * n;
  u32* off = offsets + pw * (n + 1);
  u32* cur = cursor + pw * n;
  const u32 T = blockDim.x, t = threadIdx.x;
  const u32 per = (n + T - 1) / T;
  u32 lo = t * per, hi = min(lo + per, n);
  u32 sum = 0;
  for (u32 b = lo; b < hi; b++) sum += cnt[b];
  part[t] = sum;
  __syncthreads();
  // Hillis-Steele inclusive scan of the per-thread sums
  for (u32 d = 1; d < T; d <<= 1) {
    u32 v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  u32 run = t ? part[t - 1] : 0;
  for (u32 b = lo; b < hi; b++) {
    off[b] = run;
    cur[b] = run;
    run += cnt[b];
  }
  if (t == T - 1) off[n] = part[T - 1];
}

__global__ void k_msm_scatter(MsmShape s, const u32* keys, u32* cursor, u32* sorted) {
  u64 gi = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= (u64)s.W * s.N) return;
  u32 w = (u32)(gi / s.N);
  u64 i = gi - (u64)w * s.N;
  u32 key = keys[gi];
  u32 b = key & 0x7fffffffu;
  if (!b) return;
  u32 pw = w % s.n_windows, grp = w / s.n_windows;
  u32 pos = atomicAdd(&cursor[pw * (s.nb + 1) + b], 1u);
  sorted[(u64)pw * s.n_pre * s.N + pos] = (u32)((u64)grp * s.N + i) | (key & 0x80000000u);
}

__global__ void k_msm_buckets(MsmShape s, const u32* bases, const u32* offsets,
                              const u32* sorted, u32* buckets) {
  u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (u64)s.n_windows * s.nb) return;
  u32 pw = (u32)(t / s.nb), b = (u32)(t % s.nb) + 1;
  const u32* off = offsets + pw * (s.nb + 2);
  u32 lo = off[b], hi = off[b + 1];
  const u32* list = sorted + (u64)pw * s.n_pre * s.N;
  G1Xyzz acc = G1Xyzz::identity();
  for (u32 k = lo; k < hi; k++) {
    u32 e = list[k];
    G1Aff p = load_aff(bases + 16 * (u64)(e & 0x7fffffffu));
    if (e & 0x80000000u) p = aff_neg(p);
    acc = xyzz_madd(acc, p);
  }
  store_xyzz(buckets + 32 * t, acc);
}

// segment [lo, lo + seg) of buckets (bucket index b = lo + k + 1):
// partial = sum_k (lo + k + 1) * B = T + lo * R with running sums
__global__ void k_msm_reduce(MsmShape s, const u32* buckets, u32* partials) {
  u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (u64)s.n_windows * s.nseg) return;
  u32 pw = (u32)(t / s.nseg), sg = (u32)(t % s.nseg);
  u32 lo = sg * s.seg;
  const u32* bk = buckets + 32 * ((u64)pw * s.nb + lo);
  G1Xyzz run = G1Xyzz::identity(), tot = G1Xyzz::identity();
  for (int k = (int)s.seg - 1; k >= 0; k--) {
    run = xyzz_add(run, load_xyzz(bk + 32 * k));
    tot = xyzz_add(tot, run);
  }
  if (lo) tot = xyzz_add(tot, xyzz_mul_u32(run, lo));
  store_xyzz(partials + 32 * t, tot);
}

// one block per physical window: sum of its segment partials
__global__ __launch_bounds__(256) void k_msm_sum_partials(MsmShape s, const u32* partials,
                                                          u32* wsum) {
  __shared__ u32 red[256 * 32];
  const u32 pw = blockIdx.x, t = threadIdx.x;
  G1Xyzz acc = G1Xyzz::identity();
  for (u32 k = t; k < s.nseg; k += blockDim.x)
    acc = xyzz_add(acc, load_xyzz(partials + 32 * ((u64)pw * s.nseg + k)));
  store_xyzz(red + 32 * t, acc);
  __syncthreads();
  for (u32 d = blockDim.x / 2; d > 0; d >>= 1) {
    if (t < d) {
      G1Xyzz a = load_xyzz(red + 32 * t), b = load_xyzz(red + 32 * (t + d));
      store_xyzz(red + 32 * t, xyzz_add(a, b));
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int i = 0; i < 32; i++) wsum[32 * pw + i] = red[i];
  }
}

__global__ void k_msm_final(MsmShape s, const u32* wsum, u32* result) {
  G1Xyzz acc = load_xyzz(wsum + 32 * (s.n_windows - 1));
  for (int pw = (int)s.n_windows - 2; pw >= 0; pw--) {
    for (u32 k = 0; k < s.c; k++) acc = xyzz_dbl(acc);
    acc = xyzz_add(acc, load_xyzz(wsum + 32 * pw));
  }
  store_aff(result, xyzz_to_aff(acc));
}

// table t = 2^(t * step) * table 0 (affine), t = 1 .. n_pre-1
__global__ void k_msm_precompute(u32* tables, u64 N, u32 n_pre, u32 step) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  G1Aff p = load_aff(tables + 16 * i);
  G1Xyzz acc = aff_to_xyzz(p);
  for (u32 t = 1; t < n_pre; t++) {
    for (u32 k = 0; k < step; k++) acc = xyzz_dbl(acc);
    store_aff(tables + 16 * ((u64)t * N + i), xyzz_to_aff(acc));
  }
}

// ------------------------------------------------------------------ driver
static hipError_t msm_run(const u32* bases, const u32* scalars, const MsmShape& s,
                          u32* result, char* ws, hipStream_t stream) {
  MsmLayout l = msm_layout(s);
  u32* keys = (u32*)(ws + l.keys);
  u32* sorted = (u32*)(ws + l.sorted);
  u32* counts = (u32*)(ws + l.counts);
  u32* offsets = (u32*)(ws + l.offsets);
  u32* cursor = (u32*)(ws + l.cursor);
  u32* buckets = (u32*)(ws + l.buckets);
  u32* partials = (u32*)(ws + l.partials);
  u32* wsum = (u32*)(ws + l.wsum);
  const u32 T = 256;
  u64 ncount = (u64)s.n_windows * (s.nb + 1);
  hipLaunchKernelGGL(k_zero_u32, dim3((u32)((ncount + T - 1) / T)), dim3(T), 0, stream, counts,
                     ncount);
  if (s.N) {
    { TimedScope ts_(K_MSM_DIGITS, stream); hipLaunchKernelGGL(k_msm_digits, dim3((u32)((s.N + T - 1) / T)), dim3(T), 0, stream,
                       scalars, s, keys, counts); }
  }
  { TimedScope ts_(K_MSM_SCAN, stream); hipLaunchKernelGGL(k_msm_scan, dim3(s.n_windows), dim3(1024), 0, stream, s, counts, offsets,
                     cursor); }
  u64 nk = (u64)s.W * s.N;
  if (nk) {
    { TimedScope ts_(K_MSM_SCATTER, stream); hipLaunchKernelGGL(k_msm_scatter, dim3((u32)((nk + T - 1) / T)), dim3(T), 0, stream, s,
                       keys, cursor, sorted); }
  }
  u64 nbk = (u64)s.n_windows * s.nb;
  { TimedScope ts_(K_MSM_BUCKETS, stream); hipLaunchKernelGGL(k_msm_buckets, dim3((u32)((nbk + T - 1) / T)), dim3(T), 0, stream, s, bases,
                     offsets, sorted, buckets); }
  u64 nsg = (u64)s.n_windows * s.nseg;
  { TimedScope ts_(K_MSM_REDUCE, stream); hipLaunchKernelGGL(k_msm_reduce, dim3((u32)((nsg + T - 1) / T)), dim3(T), 0, stream, s,
                     buckets, partials); }
  { TimedScope ts_(K_MSM_SUM, stream); hipLaunchKernelGGL(k_msm_sum_partials, dim3(s.n_windows), dim3(256), 0, stream, s, partials,
                     wsum); }
  { TimedScope ts_(K_MSM_FINAL, stream); hipLaunchKernelGGL(k_msm_final, dim3(1), dim3(1), 0, stream, s, wsum, result); }
  return hipGetLastError();
}

static bool valid_window(u32 c) { return c >= 4 && c <= 20; }

// Host-memory contract of the reference (msm.cu:9-41): copy, run, copy back.
static hipError_t msm_host(u32 window, u32 target, void* const* buffers,
                           unsigned long* buffer_sizes, unsigned long long len,
                           unsigned int num_cards, const unsigned int* cards,
                           unsigned int const* const* h_points, unsigned int batches,
                           unsigned int const* const* h_scaler_batch,
                           unsigned int* const* h_result) {
  IMP_REQUIRE(valid_window(window), "window");
  IMP_REQUIRE(num_cards >= 1 && cards, "cards");
  MsmShape s = msm_shape(len, window, target);
  MsmLayout l = msm_layout(s);
  // per card: points (n_pre tables), scalars, results, MSM workspace
  size_t pts = align256((size_t)s.n_pre * len * 64), scl = align256((size_t)len * 32),
         res = align256(64);
  size_t need = pts + scl + res + l.total;
  if (buffers == nullptr) {
    IMP_REQUIRE(buffer_sizes, "buffer_sizes");
    // the whole MSM runs on the first card; the others need nothing
    for (u32 i = 0; i < num_cards; i++) buffer_sizes[i] = i == 0 ? need : 0;
    return hipSuccess;
  }
  IMP_REQUIRE(h_points && (batches == 0 || (h_scaler_batch && h_result)), "null pointer");
  int prev = 0;
  IMP_CHECK(hipGetDevice(&prev));
  IMP_CHECK(hipSetDevice((int)cards[0]));
  char* buf = (char*)buffers[0];
  u32* d_pts = (u32*)buf;
  u32* d_scl = (u32*)(buf + pts);
  u32* d_res = (u32*)(buf + pts + scl);
  char* ws = buf + pts + scl + res;
  hipStream_t st;
  IMP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  for (u32 t = 0; t < s.n_pre; t++)
    IMP_CHECK(hipMemcpyAsync(d_pts + (size_t)16 * t * len, h_points[t], (size_t)len * 64,
                             hipMemcpyHostToDevice, st));
  for (u32 b = 0; b < batches; b++) {
    IMP_CHECK(hipMemcpyAsync(d_scl, h_scaler_batch[b], (size_t)len * 32, hipMemcpyHostToDevice,
                             st));
    IMP_CHECK(msm_run(d_pts, d_scl, s, d_res, ws, st));
    IMP_CHECK(hipMemcpyAsync(h_result[b], d_res, 64, hipMemcpyDeviceToHost, st));
    IMP_CHECK(hipStreamSynchronize(st));
  }
  IMP_CHECK(hipStreamSynchronize(st));
  IMP_CHECK(hipStreamDestroy(st));
  IMP_CHECK(hipSetDevice(prev));
  return hipSuccess;
}

static hipError_t msm_precompute_host(u32 window, u32 target, unsigned long long len,
                                      unsigned int* const* h_points) {
  IMP_REQUIRE(valid_window(window), "window");
  IMP_REQUIRE(h_points, "null pointer");
  MsmShape s = msm_shape(len, window, target);
  if (s.n_pre <= 1 || len == 0) return hipSuccess;
  u32* d = nullptr;
  IMP_CHECK(hipMalloc((void**)&d, (size_t)s.n_pre * len * 64));
  IMP_CHECK(hipMemcpy(d, h_points[0], (size_t)len * 64, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_msm_precompute, dim3((u32)((len + 255) / 256)), dim3(256), 0, 0, d,
                     (u64)len, s.n_pre, s.n_windows * s.c);
  IMP_CHECK(hipGetLastError());
  for (u32 t = 1; t < s.n_pre; t++)
    IMP_CHECK(hipMemcpy(h_points[t], d + (size_t)16 * t * len, (size_t)len * 64,
                        hipMemcpyDeviceToHost));
  IMP_CHECK(hipFree(d));
  return hipSuccess;
}

// default configuration of the plain `msm` symbol
static const u32 kDefaultWindow = 16;
static const u32 kDefaultTarget = 0;  // 0: one physical window per logical window

}  // namespace imp

using namespace imp;

extern "C" hipError_t msm_cfg(unsigned int window, unsigned int target_window,
                              void* const* buffers, unsigned long* buffer_sizes,
                              unsigned long long len, unsigned int, unsigned int, unsigned int,
                              unsigned int, unsigned int num_cards, const unsigned int* cards,
                              unsigned int const* const* h_points, unsigned int batches,
                              unsigned int const* const* h_scaler_batch,
                              unsigned int* const* h_result) {
  return msm_host(window, target_window, buffers, buffer_sizes, len, num_cards, cards, h_points,
                  batches, h_scaler_batch, h_result);
}

extern "C" hipError_t msm(void* const* buffers, unsigned long* buffer_sizes,
                          unsigned long long len, unsigned int batch_per_run, unsigned int parts,
                          unsigned int stage_scalers, unsigned int stage_points,
                          unsigned int num_cards, const unsigned int* cards,
                          unsigned int const* const* h_points, unsigned int batches,
                          unsigned int const* const* h_scaler_batch,
                          unsigned int* const* h_result) {
  return msm_cfg(kDefaultWindow, kDefaultTarget, buffers, buffer_sizes, len, batch_per_run,
                 parts, stage_scalers, stage_points, num_cards, cards, h_points, batches,
                 h_scaler_batch, h_result);
}

extern "C" hipError_t msm_precompute_cfg(unsigned int window, unsigned int target_window,
                                         unsigned long long len, unsigned int* const* h_points,
                                         unsigned int) {
  return msm_precompute_host(window, target_window, len, h_points);
}

extern "C" hipError_t msm_precompute(unsigned long long len, unsigned int* const* h_points,
                                     unsigned int max_cards) {
  return msm_precompute_cfg(kDefaultWindow, kDefaultTarget, len, h_points, max_cards);
}

extern "C" size_t impeller_msm_workspace_size(unsigned long long len, unsigned int window) {
  if (!valid_window(window)) return 0;
  return msm_layout(msm_shape(len, window, 0)).total;
}

extern "C" hipError_t impeller_msm(const uint32_t* bases_d, const uint32_t* scalars_d,
                                   unsigned long long len, unsigned int window,
                                   uint32_t* result_d, void* workspace, size_t workspace_bytes,
                                   hipStream_t stream) {
  IMP_REQUIRE(valid_window(window), "window");
  IMP_REQUIRE(result_d && workspace && (len == 0 || (bases_d && scalars_d)), "null pointer");
  MsmShape s = msm_shape(len, window, 0);
  IMP_REQUIRE(workspace_bytes >= msm_layout(s).total, "workspace too small");
  return msm_run(bases_d, scalars_d, s, result_d, (char*)workspace, stream);
}
