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
//   k_msm_digits      scalar -> signed digits -> (bucket key, point | sign)
//   k_sort_*          stable LSD radix sort of the keys per physical window
//                     (4-bit digits, per-tile histograms, block-local ranks)
//   k_bucket_bounds   run boundaries of each bucket in the sorted keys
//   k_bucket_tasks    each bucket split into tasks of <= T entries, one
//                     thread per task, mixed XYZZ + affine additions
//   k_bucket_combine  per-bucket partial-sum tree (groups of 2^g) so that a
//                     heavily loaded bucket never serialises on one thread
//   k_msm_reduce      sum_b b * B_b via running sums over bucket segments
//   k_msm_final       window partials -> Horner over windows -> affine
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.h"
#include "timing.h"

namespace imp {

// ------------------------------------------------------------------- shape
struct MsmShape {
  u32 c;          // window bits
  u32 W;          // logical windows = floor(254 / c) + 1
  u32 n_windows;  // physical windows (bucket sets)
  u32 n_pre;      // point tables
  u32 nb;         // buckets per physical window = 2^(c-1)
  u64 N;          // points per table
  u64 seg_len;    // sort entries per physical window = n_pre * N
  u32 tiles;      // radix-sort tiles per physical window
  u32 passes;     // radix passes (4-bit digits over c-bit keys)
  u32 T;          // entries per accumulation task
  u32 g;          // log2 of the combine group size
  u32 levels;     // combine levels
  u64 task_cap;   // task slots per physical window
  u32 seg;        // buckets per reduce segment
  u32 nseg;       // reduce segments per physical window
};

static const u32 kSortTile = 4096;  // 256 threads x 16 keys

static MsmShape msm_shape(u64 N, u32 c, u32 target_windows, u32 task = 64) {
  MsmShape s;
  s.c = c;
  s.W = 254 / c + 1;
  s.n_windows = target_windows == 0 || target_windows > s.W ? s.W : target_windows;
  s.n_pre = (s.W + s.n_windows - 1) / s.n_windows;
  s.nb = 1u << (c - 1);
  s.N = N;
  s.seg_len = (u64)s.n_pre * N;
  s.tiles = (u32)((s.seg_len + kSortTile - 1) / kSortTile);
  if (s.tiles == 0) s.tiles = 1;
  s.passes = (c + 3) / 4;
  s.T = task;
  s.g = 5;
  s.task_cap = s.seg_len / s.T + s.nb + 1;
  u64 max_tasks = s.seg_len / s.T + 1;
  s.levels = 0;
  while ((1ull << (s.g * s.levels)) < max_tasks && s.levels < 6) s.levels++;
  s.seg = s.nb >= 8 ? 8 : s.nb;
  s.nseg = s.nb / s.seg;
  return s;
}

struct MsmLayout {
  size_t keysA, valsA, keysB, valsB, hist, bstart, bend, ntask, toff, partA, partB, rparts, wsum,
      total;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static MsmLayout msm_layout(const MsmShape& s) {
  MsmLayout l;
  size_t o = 0;
  const size_t ent = (size_t)s.n_windows * s.seg_len * 4;
  l.keysA = o; o = align256(o + ent);
  l.valsA = o; o = align256(o + ent);
  l.keysB = o; o = align256(o + ent);
  l.valsB = o; o = align256(o + ent);
  l.hist = o; o = align256(o + (size_t)s.n_windows * 16 * s.tiles * 4 + 4 * s.n_windows);
  l.bstart = o; o = align256(o + (size_t)s.n_windows * (s.nb + 1) * 4);
  l.bend = o; o = align256(o + (size_t)s.n_windows * (s.nb + 1) * 4);
  l.ntask = o; o = align256(o + (size_t)s.n_windows * (s.nb + 1) * 4);
  l.toff = o; o = align256(o + (size_t)s.n_windows * (s.nb + 2) * 4);
  l.partA = o; o = align256(o + (size_t)s.n_windows * s.task_cap * 128);
  l.partB = o; o = align256(o + (size_t)s.n_windows * s.task_cap * 128);
  l.rparts = o; o = align256(o + (size_t)s.n_windows * s.nseg * 128);
  l.wsum = o; o = align256(o + (size_t)s.n_windows * 128);
  l.total = o;
  return l;
}

// ------------------------------------------------------------------ kernels
__global__ void k_zero_u32(u32* p, u64 n) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Signed c-bit digits (msm_impl.cuh:22-33; a digit equal to 2^(c-1) stays
// positive, so the top window never carries out).  Logical window w goes to
// physical window w % n_windows at table offset (w / n_windows) * N
// (msm_impl.cuh:56-67).  key = bucket |digit|, val = point index | sign << 31.
__global__ void k_msm_digits(const u32* scalars, MsmShape s, u32* keys, u32* vals) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.N) return;
  Fr v = to_number<FrParams>(load_fe<Fr>(scalars + 8 * i));
  const u32 c = s.c, half = 1u << (c - 1), mask = (1u << c) - 1;
  u32 carry = 0;
  const u32 slots = s.n_windows * s.n_pre;
  for (u32 w = 0; w < slots; w++) {
    u32 pw = w % s.n_windows, grp = w / s.n_windows;
    u64 idx = (u64)pw * s.seg_len + (u64)grp * s.N + i;
    u32 key = 0, sign = 0;
    if (w < s.W) {
      u32 bit = w * c, limb = bit >> 5, off = bit & 31;
      u64 word = v.v[limb];
      if (limb + 1 < 8) word |= (u64)v.v[limb + 1] << 32;
      u32 d = (u32)((word >> off) & mask) + carry;
      if (d > half) {
        key = (1u << c) - d;
        sign = 0x80000000u;
        carry = 1;
      } else {
        key = d;
        carry = 0;
      }
    }
    keys[idx] = key;
    vals[idx] = (u32)((u64)grp * s.N + i) | sign;
  }
}

// ---------------------------------------------------- LSD radix sort (4-bit)
// Stable per pass: tiles of 16 consecutive keys per thread (blocked), block
// ranks in (digit, thread) order, global bases from the scanned per-tile
// histograms hist[(pw * 16 + digit) * tiles + tile].
__global__ __launch_bounds__(256) void k_sort_hist(const u32* keys, MsmShape s, u32 shift,
                                                   u32* hist) {
  __shared__ u32 cnt[16];
  const u32 tile = blockIdx.x, pw = blockIdx.y, t = threadIdx.x;
  if (t < 16) cnt[t] = 0;
  __syncthreads();
  const u32* kk = keys + (u64)pw * s.seg_len;
  u64 base = (u64)tile * kSortTile + (u64)t * 16;
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    u64 idx = base + j;
    if (idx < s.seg_len) atomicAdd(&cnt[(kk[idx] >> shift) & 15u], 1u);
  }
  __syncthreads();
  if (t < 16) hist[((u64)pw * 16 + t) * s.tiles + tile] = cnt[t];
}

// exclusive scan of len entries per segment (in place allowed); with
// write_total, out[len] = total (out_stride must then exceed len)
__global__ __launch_bounds__(1024) void k_seg_scan(const u32* in, u32* out, u64 len,
                                                   u64 in_stride, u64 out_stride,
                                                   int write_total) {
  __shared__ u32 part[1024];
  const u32 T = blockDim.x, t = threadIdx.x;
  const u32* src = in + (u64)blockIdx.x * in_stride;
  u32* dst = out + (u64)blockIdx.x * out_stride;
  const u64 per = (len + T - 1) / T;
  u64 lo = (u64)t * per, hi = lo + per < len ? lo + per : len;
  u32 sum = 0;
  for (u64 b = lo; b < hi; b++) sum += src[b];
  part[t] = sum;
  __syncthreads();
  for (u32 d = 1; d < T; d <<= 1) {
    u32 v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  u32 run = t ? part[t - 1] : 0;
  for (u64 b = lo; b < hi; b++) {
    u32 x = src[b];
    dst[b] = run;
    run += x;
  }
  if (write_total && t == T - 1) dst[len] = part[T - 1];
}

__global__ __launch_bounds__(256) void k_sort_scatter(const u32* ksrc, const u32* vsrc, u32* kdst,
                                                      u32* vdst, MsmShape s, u32 shift,
                                                      const u32* hist) {
  __shared__ u32 cnt[16 * 256];
  __shared__ u32 part[256];
  __shared__ u32 gbase[16];
  __shared__ u32 dbase[16];
  const u32 tile = blockIdx.x, pw = blockIdx.y, t = threadIdx.x;
  const u64 seg0 = (u64)pw * s.seg_len;
  const u64 base = (u64)tile * kSortTile + (u64)t * 16;
  u32 key[16], val[16];
#pragma unroll
  for (u32 d = 0; d < 16; d++) cnt[d * 256 + t] = 0;
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    u64 idx = base + j;
    if (idx < s.seg_len) {
      key[j] = ksrc[seg0 + idx];
      val[j] = vsrc[seg0 + idx];
      cnt[((key[j] >> shift) & 15u) * 256 + t] += 1;
    }
  }
  __syncthreads();
  // exclusive scan of cnt in (digit, thread) order: thread t owns [16t, 16t+16)
  u32 loc[16];
  u32 sum = 0;
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    loc[j] = cnt[t * 16 + j];
    sum += loc[j];
  }
  part[t] = sum;
  __syncthreads();
  for (u32 d = 1; d < 256; d <<= 1) {
    u32 v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  u32 run = t ? part[t - 1] : 0;
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    cnt[t * 16 + j] = run;
    run += loc[j];
  }
  __syncthreads();
  if (t < 16) {
    gbase[t] = hist[((u64)pw * 16 + t) * s.tiles + tile];
    dbase[t] = cnt[t * 256];
  }
  __syncthreads();
  // rank within digit d = scan(d, t) - scan(d, 0); cnt column t is private
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    u64 idx = base + j;
    if (idx < s.seg_len) {
      u32 d = (key[j] >> shift) & 15u;
      u32 r = cnt[d * 256 + t];
      cnt[d * 256 + t] = r + 1;
      u64 pos = (u64)gbase[d] + r - dbase[d];
      kdst[seg0 + pos] = key[j];
      vdst[seg0 + pos] = val[j];
    }
  }
}

// bucket run bounds in the sorted keys (bucket 0 = zero digit, skipped)
__global__ void k_bucket_bounds(const u32* keys, MsmShape s, u32* bstart, u32* bend) {
  u64 gi = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= (u64)s.n_windows * s.seg_len) return;
  u32 pw = (u32)(gi / s.seg_len);
  u64 i = gi - (u64)pw * s.seg_len;
  const u32* kk = keys + (u64)pw * s.seg_len;
  u32 k = kk[i];
  if (k == 0) return;
  u32* st = bstart + (u64)pw * (s.nb + 1);
  u32* en = bend + (u64)pw * (s.nb + 1);
  if (i == 0 || kk[i - 1] != k) st[k] = (u32)i;
  if (i + 1 == s.seg_len || kk[i + 1] != k) en[k] = (u32)(i + 1);
}

__global__ void k_task_counts(MsmShape s, const u32* bstart, const u32* bend, u32* ntask) {
  u64 gi = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= (u64)s.n_windows * (s.nb + 1)) return;
  u32 b = (u32)(gi % (s.nb + 1));
  u32 cnt = bend[gi] - bstart[gi];
  ntask[gi] = b == 0 ? 0u : (cnt + s.T - 1) / s.T;
}

// largest b in [1, nb] with toff[b] <= t
__device__ __forceinline__ u32 find_bucket(const u32* toff, u32 nb, u32 t) {
  u32 lo = 1, hi = nb;
  while (lo < hi) {
    u32 mid = (lo + hi + 1) >> 1;
    if (toff[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// one task = up to T consecutive entries of one bucket, mixed XYZZ + affine
// additions; the next point is loaded while the current one is added.
__global__ __launch_bounds__(256) void k_bucket_tasks(const u32* bases, const u32* vals,
                                                      MsmShape s, const u32* bstart,
                                                      const u32* bend, const u32* toff,
                                                      u32* parts) {
  u64 gi = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  u32 pw = (u32)(gi / s.task_cap);
  if (pw >= s.n_windows) return;
  u32 t = (u32)(gi - (u64)pw * s.task_cap);
  const u32* to = toff + (u64)pw * (s.nb + 2);
  if (t >= to[s.nb + 1]) return;
  u32 b = find_bucket(to, s.nb, t);
  u32 k = t - to[b];
  u32 lo = bstart[(u64)pw * (s.nb + 1) + b] + k * s.T;
  u32 be = bend[(u64)pw * (s.nb + 1) + b];
  u32 hi = lo + s.T < be ? lo + s.T : be;
  const u32* vv = vals + (u64)pw * s.seg_len;
  G1Xyzz acc = G1Xyzz::identity();
  u32 e = vv[lo];
  G1Aff p = load_aff(bases + 16 * (u64)(e & 0x7fffffffu));
  for (u32 j = lo; j < hi; j++) {
    G1Aff cur = p;
    u32 ce = e;
    if (j + 1 < hi) {
      e = vv[j + 1];
      p = load_aff(bases + 16 * (u64)(e & 0x7fffffffu));
    }
    if (ce & 0x80000000u) cur = aff_neg(cur);
    acc = xyzz_madd(acc, cur);
  }
  store_xyzz(parts + 32 * ((u64)pw * s.task_cap + t), acc);
}

// level l of the per-bucket partial-sum tree: group k of bucket b sums
// src[toff_b + kG .. toff_b + kG + G) (live entries only) into dst[toff_b + k]
__global__ __launch_bounds__(256) void k_bucket_combine(const u32* src, u32* dst, MsmShape s,
                                                        const u32* toff, u32 level) {
  u64 gi = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  u32 pw = (u32)(gi / s.task_cap);
  if (pw >= s.n_windows) return;
  u32 t = (u32)(gi - (u64)pw * s.task_cap);
  const u32* to = toff + (u64)pw * (s.nb + 2);
  if (t >= to[s.nb + 1]) return;
  u32 b = find_bucket(to, s.nb, t);
  u32 k = t - to[b];
  u32 nt = to[b + 1] - to[b];
  u32 sh_prev = s.g * (level - 1), sh = s.g * level;
  u32 n_prev = (u32)(((u64)nt + (1ull << sh_prev) - 1) >> sh_prev);
  u32 n_cur = (u32)(((u64)nt + (1ull << sh) - 1) >> sh);
  if (k >= n_cur) return;
  const u64 base = (u64)pw * s.task_cap + to[b];
  const u32 G = 1u << s.g;
  G1Xyzz acc = load_xyzz(src + 32 * (base + (u64)k * G));
  for (u32 j = 1; j < G && k * G + j < n_prev; j++)
    acc = xyzz_add(acc, load_xyzz(src + 32 * (base + (u64)k * G + j)));
  store_xyzz(dst + 32 * (base + k), acc);
}

// segment [lo, lo + seg) of buckets (bucket number b = lo + k + 1):
// partial = sum_k (lo + k + 1) * B = T + lo * R with running sums
__global__ void k_msm_reduce(MsmShape s, const u32* parts, const u32* toff, u32* rparts) {
  u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (u64)s.n_windows * s.nseg) return;
  u32 pw = (u32)(t / s.nseg), sg = (u32)(t % s.nseg);
  u32 lo = sg * s.seg;
  const u32* to = toff + (u64)pw * (s.nb + 2);
  const u32* pp = parts + 32 * (u64)pw * s.task_cap;
  G1Xyzz run = G1Xyzz::identity(), tot = G1Xyzz::identity();
  for (int k = (int)s.seg - 1; k >= 0; k--) {
    u32 b = lo + k + 1;
    if (to[b + 1] > to[b]) run = xyzz_add(run, load_xyzz(pp + 32 * (u64)to[b]));
    tot = xyzz_add(tot, run);
  }
  if (lo) tot = xyzz_add(tot, xyzz_mul_u32(run, lo));
  store_xyzz(rparts + 32 * t, tot);
}

// one block per physical window: sum of its segment partials
__global__ __launch_bounds__(256) void k_msm_sum_partials(MsmShape s, const u32* rparts,
                                                          u32* wsum) {
  __shared__ u32 red[256 * 32];
  const u32 pw = blockIdx.x, t = threadIdx.x;
  G1Xyzz acc = G1Xyzz::identity();
  for (u32 k = t; k < s.nseg; k += blockDim.x)
    acc = xyzz_add(acc, load_xyzz(rparts + 32 * ((u64)pw * s.nseg + k)));
  store_xyzz(red + 32 * t, acc);
  __syncthreads();
  for (u32 d = blockDim.x / 2; d > 0; d >>= 1) {
    if (t < d) {
      G1Xyzz a = load_xyzz(red + 32 * t), b = load_xyzz(red + 32 * (t + d));
      store_xyzz(red + 32 * t, xyzz_add(a, b));
    }
    __syncthreads();
  }
  if (t < 32) wsum[32 * pw + t] = red[t];
}

// Horner over physical windows, then affine
__global__ void k_msm_final(MsmShape s, const u32* wsum, u32* result) {
  G1Xyzz acc = load_xyzz(wsum + 32 * (s.n_windows - 1));
  for (int pw = (int)s.n_windows - 2; pw >= 0; pw--) {
    for (u32 k = 0; k < s.c; k++) acc = xyzz_dbl(acc);
    acc = xyzz_add(acc, load_xyzz(wsum + 32 * pw));
  }
  store_aff(result, xyzz_to_aff(acc));
}

__global__ void k_store_identity_aff(u32* result) {
  if (threadIdx.x < 16) result[threadIdx.x] = 0;
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
static inline dim3 grid1(u64 n, u32 block) { return dim3((u32)((n + block - 1) / block)); }

// bases: n_pre tables of N affine points (table t = 2^(t * n_windows * c) * P)
static hipError_t msm_run(const u32* bases, const u32* scalars, const MsmShape& s,
                          u32* result, char* ws, hipStream_t stream) {
  if (s.N == 0) {
    hipLaunchKernelGGL(k_store_identity_aff, dim3(1), dim3(64), 0, stream, result);
    return hipGetLastError();
  }
  MsmLayout l = msm_layout(s);
  u32* kA = (u32*)(ws + l.keysA);
  u32* vA = (u32*)(ws + l.valsA);
  u32* kB = (u32*)(ws + l.keysB);
  u32* vB = (u32*)(ws + l.valsB);
  u32* hist = (u32*)(ws + l.hist);
  u32* bstart = (u32*)(ws + l.bstart);
  u32* bend = (u32*)(ws + l.bend);
  u32* ntask = (u32*)(ws + l.ntask);
  u32* toff = (u32*)(ws + l.toff);
  u32* partA = (u32*)(ws + l.partA);
  u32* partB = (u32*)(ws + l.partB);
  u32* rparts = (u32*)(ws + l.rparts);
  u32* wsum = (u32*)(ws + l.wsum);
  const u32 TB = 256;
  {
    TimedScope ts_(K_MSM_DIGITS, stream);
    hipLaunchKernelGGL(k_msm_digits, grid1(s.N, TB), dim3(TB), 0, stream, scalars, s, kA, vA);
  }
  {
    TimedScope ts_(K_MSM_SORT, stream);
    u32 *ks = kA, *vs = vA, *kd = kB, *vd = vB;
    dim3 tg(s.tiles, s.n_windows);
    for (u32 p = 0; p < s.passes; p++) {
      hipLaunchKernelGGL(k_sort_hist, tg, dim3(256), 0, stream, ks, s, 4 * p, hist);
      hipLaunchKernelGGL(k_seg_scan, dim3(s.n_windows), dim3(1024), 0, stream, hist, hist,
                         (u64)16 * s.tiles, (u64)16 * s.tiles, (u64)16 * s.tiles, 0);
      hipLaunchKernelGGL(k_sort_scatter, tg, dim3(256), 0, stream, ks, vs, kd, vd, s, 4 * p,
                         hist);
      u32* tk = ks; ks = kd; kd = tk;
      u32* tv = vs; vs = vd; vd = tv;
    }
    kA = ks;  // sorted keys / values
    vA = vs;
  }
  {
    TimedScope ts_(K_MSM_SCAN, stream);
    u64 nbw = (u64)s.n_windows * (s.nb + 1);
    hipLaunchKernelGGL(k_zero_u32, grid1(nbw, TB), dim3(TB), 0, stream, bstart, nbw);
    hipLaunchKernelGGL(k_zero_u32, grid1(nbw, TB), dim3(TB), 0, stream, bend, nbw);
    hipLaunchKernelGGL(k_bucket_bounds, grid1((u64)s.n_windows * s.seg_len, TB), dim3(TB), 0,
                       stream, kA, s, bstart, bend);
    hipLaunchKernelGGL(k_task_counts, grid1(nbw, TB), dim3(TB), 0, stream, s, bstart, bend,
                       ntask);
    hipLaunchKernelGGL(k_seg_scan, dim3(s.n_windows), dim3(1024), 0, stream, ntask, toff,
                       (u64)(s.nb + 1), (u64)(s.nb + 1), (u64)(s.nb + 2), 1);
  }
  const u64 ntasks = (u64)s.n_windows * s.task_cap;
  {
    TimedScope ts_(K_MSM_BUCKETS, stream);
    hipLaunchKernelGGL(k_bucket_tasks, grid1(ntasks, TB), dim3(TB), 0, stream, bases, vA, s,
                       bstart, bend, toff, partA);
  }
  u32* fin = partA;
  {
    TimedScope ts_(K_MSM_COMBINE, stream);
    u32 *src = partA, *dst = partB;
    for (u32 lv = 1; lv <= s.levels; lv++) {
      hipLaunchKernelGGL(k_bucket_combine, grid1(ntasks, TB), dim3(TB), 0, stream, src, dst, s,
                         toff, lv);
      u32* tmp = src; src = dst; dst = tmp;
    }
    fin = src;
  }
  {
    TimedScope ts_(K_MSM_REDUCE, stream);
    hipLaunchKernelGGL(k_msm_reduce, grid1((u64)s.n_windows * s.nseg, TB), dim3(TB), 0, stream,
                       s, fin, toff, rparts);
  }
  {
    TimedScope ts_(K_MSM_SUM, stream);
    hipLaunchKernelGGL(k_msm_sum_partials, dim3(s.n_windows), dim3(256), 0, stream, s, rparts,
                       wsum);
  }
  {
    TimedScope ts_(K_MSM_FINAL, stream);
    hipLaunchKernelGGL(k_msm_final, dim3(1), dim3(1), 0, stream, s, wsum, result);
  }
  return hipGetLastError();
}

static bool valid_window(u32 c) { return c >= 4 && c <= 20; }

static hipError_t precompute_device(u32* tables, u64 N, const MsmShape& s, hipStream_t stream) {
  if (s.n_pre <= 1 || N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_msm_precompute, grid1(N, 128), dim3(128), 0, stream, tables, N, s.n_pre,
                     s.n_windows * s.c);
  return hipGetLastError();
}

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
  IMP_CHECK(precompute_device(d, len, s, 0));
  IMP_CHECK(hipDeviceSynchronize());
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

extern "C" size_t impeller_msm_cfg_workspace_size(unsigned long long len, unsigned int window,
                                                  unsigned int target_windows,
                                                  unsigned int task) {
  if (!valid_window(window) || task == 0) return 0;
  return msm_layout(msm_shape(len, window, target_windows, task)).total;
}

extern "C" unsigned int impeller_msm_tables(unsigned int window, unsigned int target_windows) {
  if (!valid_window(window)) return 0;
  return msm_shape(1, window, target_windows).n_pre;
}

extern "C" hipError_t impeller_msm_precompute(uint32_t* tables_d, unsigned long long len,
                                              unsigned int window, unsigned int target_windows,
                                              hipStream_t stream) {
  IMP_REQUIRE(valid_window(window), "window");
  IMP_REQUIRE(tables_d || len == 0, "null pointer");
  return precompute_device(tables_d, len, msm_shape(len, window, target_windows), stream);
}

extern "C" hipError_t impeller_msm_cfg(const uint32_t* tables_d, const uint32_t* scalars_d,
                                       unsigned long long len, unsigned int window,
                                       unsigned int target_windows, unsigned int task,
                                       uint32_t* result_d, void* workspace,
                                       size_t workspace_bytes, hipStream_t stream) {
  IMP_REQUIRE(valid_window(window) && task > 0 && task <= 4096, "window / task");
  IMP_REQUIRE(result_d && workspace && (len == 0 || (tables_d && scalars_d)), "null pointer");
  MsmShape s = msm_shape(len, window, target_windows, task);
  IMP_REQUIRE(workspace_bytes >= msm_layout(s).total, "workspace too small");
  return msm_run(tables_d, scalars_d, s, result_d, (char*)workspace, stream);
}
