// ntt.hip -- in-place BN254 Fr NTT for gfx950.
//
// Transform: the radix-2 decimation-in-time network of halo2 best_fft
// (core/tests/msm_halo2.rs:221-324), which the reference's SSIP NTT matches
// bit for bit even for an omega that is not a root of unity
// (core/tests/ntt.rs:63,151-153).  For that to hold, every butterfly here is
// the same as in best_fft: after the bit-reversal permutation, level l pairs
// positions p and p + 2^l (bit l of p clear) with twiddle
// omega^((p mod 2^l) * 2^(L-l-1)), a table entry below n/2, and computes
// (a + w*b, a - w*b).  Only the grouping of levels into HBM passes differs.
//
// Passes (all in place, no workspace):
//   1. bit-reversal permutation: tile pairs (b, rev b) of 32 x 32 elements
//      swapped through LDS, optionally multiplying x[q] by shift^q (coset);
//   2. ceil(L / 8) DIT passes; each work-group stages a tile of G = 2^D
//      positions x C columns in LDS and runs D consecutive levels, columns
//      chosen so that every global access is a run of >= 4 consecutive
//      32-byte elements (one 128-byte line);
//   3. the last pass can multiply each output by post(i) (n^-1 and/or
//      shift^-i for inverse and inverse-coset transforms).
#include <mutex>
#include <vector>

#include "common.h"
#include "timing.h"

namespace imp {

// --------------------------------------------------------- on-the-fly factors
// factor(i) = T0[i & 1023] * T1[(i >> 10) & 1023] * T2[i >> 20]; digits the
// transform size never reaches are skipped.  The tables hold exact powers, so
// the product equals the direct power bit for bit.
struct Factor {
  const u32* tab;  // 3 * 1024 elements, or nullptr (factor 1)
  u32 ndig;        // number of 10-bit digits used (1..3)
};

__device__ __forceinline__ Fr factor_at(const Factor& f, u64 i) {
  Fr r = load_fe<Fr>(f.tab + 8 * (i & 1023));
  if (f.ndig > 1) r = mul<FrParams>(r, load_fe<Fr>(f.tab + 8 * (1024 + ((i >> 10) & 1023))));
  if (f.ndig > 2) r = mul<FrParams>(r, load_fe<Fr>(f.tab + 8 * (2048 + ((i >> 20) & 1023))));
  return r;
}

// tab[d * 1024 + j] = scale^(d==0) * base^(j * 1024^d)
__global__ void k_factor_tables(u32* tab, const u32* base_p, const u32* scale_p, u32 ndig) {
  u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 1024 * ndig) return;
  u32 d = t >> 10, j = t & 1023;
  Fr base = base_p ? load_fe<Fr>(base_p) : Fr::one();
  for (u32 k = 0; k < 10 * d; k++) base = sqr<FrParams>(base);
  Fr r = pow_u32<FrParams>(base, j);
  if (d == 0 && scale_p) r = mul<FrParams>(r, load_fe<Fr>(scale_p));
  store_fe<Fr>(tab + 8 * t, r);
}

// --------------------------------------------------------------- twiddles
struct Twiddle {
  const u32* tab;  // omega^e for e < n/2 (direct mode)
  const u32* lo;   // two-level mode: omega^e = lo[e & mask] * hi[e >> h]
  const u32* hi;
  u32 h;
};

template <bool TWO>
__device__ __forceinline__ Fr twiddle_at(const Twiddle& tw, u64 e) {
  if (!TWO) return load_fe<Fr>(tw.tab + 8 * e);
  u64 mask = (1ull << tw.h) - 1;
  Fr a = load_fe<Fr>(tw.lo + 8 * (e & mask));
  u64 eh = e >> tw.h;
  if (eh == 0) return a;
  return mul<FrParams>(a, load_fe<Fr>(tw.hi + 8 * eh));
}

// --------------------------------------------------------- bit reversal
// q = (a: 5 bits | b: L-10 bits | c: 5 bits), rev(q) = (rev c | rev b | rev a)
template <bool PRO>
__global__ __launch_bounds__(256) void k_bitrev_tiles(View x, u32 L, Factor pro) {
  extern __shared__ uint4 lds4[];
  u32* t0 = reinterpret_cast<u32*>(lds4);
  u32* t1 = t0 + 1024 * 8;
  const u32 mb = L - 10;
  const u64 b = blockIdx.x;
  const u64 rb = bit_reverse(b, mb);
  if (b > rb) return;
  const bool self = (b == rb);
  for (u32 idx = threadIdx.x; idx < 1024; idx += blockDim.x) {
    u64 a = idx >> 5, c = idx & 31;
    u64 q0 = (a << (L - 5)) | (b << 5) | c;
    Fr v0 = load_fe<Fr>(x.base + 8 * view_phys(x, q0));
    if (PRO) v0 = mul<FrParams>(v0, factor_at(pro, q0));
    store_fe<Fr>(t0 + 8 * idx, v0);
    if (!self) {
      u64 q1 = (a << (L - 5)) | (rb << 5) | c;
      Fr v1 = load_fe<Fr>(x.base + 8 * view_phys(x, q1));
      if (PRO) v1 = mul<FrParams>(v1, factor_at(pro, q1));
      store_fe<Fr>(t1 + 8 * idx, v1);
    }
  }
  __syncthreads();
  for (u32 idx = threadIdx.x; idx < 1024; idx += blockDim.x) {
    u32 a = idx >> 5, c = idx & 31;
    u32 src = (brev_bits(c, 5) << 5) | brev_bits(a, 5);
    u64 q0 = ((u64)a << (L - 5)) | (b << 5) | c;
    const u32* from0 = self ? t0 : t1;
    store_fe<Fr>(x.base + 8 * view_phys(x, q0), load_fe<Fr>(from0 + 8 * src));
    if (!self) {
      u64 q1 = ((u64)a << (L - 5)) | (rb << 5) | c;
      store_fe<Fr>(x.base + 8 * view_phys(x, q1), load_fe<Fr>(t0 + 8 * src));
    }
  }
}

// whole array (n <= 512) in one work-group
template <bool PRO>
__global__ __launch_bounds__(256) void k_bitrev_small(View x, u32 L, Factor pro) {
  extern __shared__ uint4 lds4[];
  u32* t = reinterpret_cast<u32*>(lds4);
  const u32 n = 1u << L;
  for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
    Fr v = load_fe<Fr>(x.base + 8 * view_phys(x, i));
    if (PRO) v = mul<FrParams>(v, factor_at(pro, i));
    store_fe<Fr>(t + 8 * i, v);
  }
  __syncthreads();
  for (u32 i = threadIdx.x; i < n; i += blockDim.x)
    store_fe<Fr>(x.base + 8 * view_phys(x, i), load_fe<Fr>(t + 8 * brev_bits(i, L)));
}

// ------------------------------------------------------------- DIT pass
// Levels s .. s+D-1 on positions p = hi*2^(s+D) + g*2^s + lo.
// contiguous (s == 0): work-group = C consecutive groups of G positions.
// strided (2^s >= C): work-group = fixed hi, C consecutive lo.
template <bool TWO, bool EPI>
__global__ __launch_bounds__(256) void k_dit_pass(View x, u32 L, u32 s, u32 D, u32 logC,
                                                  Twiddle tw, Factor epi) {
  extern __shared__ uint4 lds4[];
  u32* t = reinterpret_cast<u32*>(lds4);
  const u32 C = 1u << logC, G = 1u << D, GC = G * C;
  const u64 wg = blockIdx.x;
  const bool contig = (s == 0);
  u64 hi = 0, lo0 = 0;
  if (!contig) {
    hi = wg >> (s - logC);
    lo0 = (wg & ((1ull << (s - logC)) - 1)) << logC;
  }

  // load: consecutive threads read consecutive addresses
  for (u32 idx = threadIdx.x; idx < GC; idx += blockDim.x) {
    u32 g, c;
    u64 p;
    if (contig) {
      c = idx >> D;
      g = idx & (G - 1);
      p = (((wg << logC) + c) << D) | g;
    } else {
      g = idx >> logC;
      c = idx & (C - 1);
      p = (hi << (s + D)) | ((u64)g << s) | (lo0 + c);
    }
    store_fe<Fr>(t + 8 * (g * C + c), load_fe<Fr>(x.base + 8 * view_phys(x, p)));
  }
  __syncthreads();

  const u32 nb = GC >> 1;
  for (u32 j = 0; j < D; j++) {
    const u32 l = s + j;
    const u32 sh = L - l - 1;
    for (u32 bi = threadIdx.x; bi < nb; bi += blockDim.x) {
      u32 c = bi & (C - 1), k = bi >> logC;
      u32 glo = k & ((1u << j) - 1);
      u32 g0 = ((k >> j) << (j + 1)) | glo, g1 = g0 | (1u << j);
      u64 lo = contig ? 0 : lo0 + c;
      u64 e = ((((u64)glo) << s) + lo) << sh;
      u32* pa = t + 8 * (g0 * C + c);
      u32* pb = t + 8 * (g1 * C + c);
      Fr a = load_fe<Fr>(pa), bv = load_fe<Fr>(pb);
      if (e != 0) bv = mul<FrParams>(bv, twiddle_at<TWO>(tw, e));
      store_fe<Fr>(pa, add<FrParams>(a, bv));
      store_fe<Fr>(pb, sub<FrParams>(a, bv));
    }
    __syncthreads();
  }

  for (u32 idx = threadIdx.x; idx < GC; idx += blockDim.x) {
    u32 g, c;
    u64 p;
    if (contig) {
      c = idx >> D;
      g = idx & (G - 1);
      p = (((wg << logC) + c) << D) | g;
    } else {
      g = idx >> logC;
      c = idx & (C - 1);
      p = (hi << (s + D)) | ((u64)g << s) | (lo0 + c);
    }
    Fr v = load_fe<Fr>(t + 8 * (g * C + c));
    if (EPI) v = mul<FrParams>(v, factor_at(epi, p));
    store_fe<Fr>(x.base + 8 * view_phys(x, p), v);
  }
}

// ------------------------------------------------------------ elementwise
// distribute_powers.cuh:5-10 semantics over a PolyPtr view.
__global__ void k_distribute_powers(View x, u64 len, const u32* powers, u64 power_num,
                                    PolySlice sl) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  u64 origin = (sl.begin + i) % sl.len;
  u32* p = x.base + 8 * view_phys(x, i);
  Fr v = mul<FrParams>(load_fe<Fr>(p), load_fe<Fr>(powers + 8 * (origin % power_num)));
  store_fe<Fr>(p, v);
}

// out[i] = base^i for i < len: each thread runs a chunk of sequential products
__global__ void k_powers(u32* out, u64 len, const u32* base_p, u32 chunk) {
  u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  u64 start = t * chunk;
  if (start >= len) return;
  Fr base = load_fe<Fr>(base_p);
  // base^start by binary powering over the 64-bit exponent
  Fr acc = Fr::one(), sq = base;
  for (u64 e = start; e; e >>= 1) {
    if (e & 1) acc = mul<FrParams>(acc, sq);
    sq = sqr<FrParams>(sq);
  }
  u64 end = start + chunk < len ? start + chunk : len;
  for (u64 i = start; i < end; i++) {
    store_fe<Fr>(out + 8 * i, acc);
    acc = mul<FrParams>(acc, base);
  }
}

// omega^e for e < count from omegas[k] = omega^(2^k): product over set bits
__global__ void k_pow_from_omegas(u32* out, u64 count, u64 stride_log, const u32* omegas) {
  u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  u64 e = j << stride_log;
  Fr acc = Fr::one();
  for (u32 k = 0; e && k < 32; k++, e >>= 1)
    if (e & 1) acc = mul<FrParams>(acc, load_fe<Fr>(omegas + 8 * k));
  store_fe<Fr>(out + 8 * j, acc);
}

// om[i] = om[i-1]^2 for i < count (single thread)
__global__ void k_square_chain(u32* om, u32 count) {
  Fr w = load_fe<Fr>(om);
  for (u32 i = 1; i < count; i++) {
    w = sqr<FrParams>(w);
    store_fe<Fr>(om + 8 * i, w);
  }
}

// n^-1 for n = 2^L: Fermat inversion of from_number(2^L)
__global__ void k_ninv(u32* out, u32 L) {
  Fr n = Fr::zero();
  n.v[L / 32] = 1u << (L % 32);
  n = from_number<FrParams>(n);
  store_fe<Fr>(out, inv<FrParams>(n));
}

// ----------------------------------------------------------------- planner
struct PassPlan {
  u32 n;
  u32 d[32];
};

static void default_plan(u32 L, PassPlan* p) {
  // passes of at most 8 levels, as equal as possible
  u32 np = (L + 7) / 8;
  if (np == 0) np = 1;
  p->n = np;
  for (u32 i = 0; i < np; i++) p->d[i] = L / np + (i < L % np ? 1 : 0);
}

static u32 ilog2(u64 x) {
  u32 r = 0;
  while ((1ull << (r + 1)) <= x) r++;
  return r;
}

struct NttArgs {
  View x;
  u32 L;
  Twiddle tw;
  bool two_level;
  Factor pro;  // tab == nullptr: none
  Factor epi;
  const PassPlan* plan;
};

static hipError_t run_ntt(const NttArgs& a, hipStream_t stream) {
  const u32 L = a.L;
  // 1. bit reversal (+ prologue)
  {
  TimedScope ts(K_NTT_BITREV, stream);
  if (L >= 10) {
    u64 grid = 1ull << (L - 10);
    size_t lds = 2 * 1024 * 32;
    if (a.pro.tab)
      hipLaunchKernelGGL((k_bitrev_tiles<true>), dim3((u32)grid), dim3(256), lds, stream, a.x, L,
                         a.pro);
    else
      hipLaunchKernelGGL((k_bitrev_tiles<false>), dim3((u32)grid), dim3(256), lds, stream, a.x,
                         L, a.pro);
  } else {
    size_t lds = (size_t)32 << L;
    if (a.pro.tab)
      hipLaunchKernelGGL((k_bitrev_small<true>), dim3(1), dim3(256), lds, stream, a.x, L, a.pro);
    else
      hipLaunchKernelGGL((k_bitrev_small<false>), dim3(1), dim3(256), lds, stream, a.x, L,
                         a.pro);
  }
  }
  IMP_CHECK(hipGetLastError());
  if (L == 0) return hipSuccess;

  // 2. DIT passes
  u32 s = 0;
  for (u32 pi = 0; pi < a.plan->n; pi++) {
    u32 D = a.plan->d[pi];
    if (D == 0) continue;
    bool last = (pi + 1 == a.plan->n);
    u32 logC;
    u64 grid;
    // at most 1024 elements (32 KiB) staged per work-group
    u32 maxLogC = D >= 10 ? 0 : 10 - D;
    if (s == 0) {
      u32 groups_log = L - D;
      logC = maxLogC < groups_log ? maxLogC : groups_log;
      grid = 1ull << (L - D - logC);
    } else {
      logC = maxLogC < s ? maxLogC : s;
      grid = 1ull << (L - D - logC);
    }
    size_t lds = ((size_t)32 << D) << logC;
    bool epi = last && a.epi.tab;
    TimedScope ts(K_NTT_DIT, stream);
#define IMP_LAUNCH_DIT(TWO, EPI)                                                              \
  hipLaunchKernelGGL((k_dit_pass<TWO, EPI>), dim3((u32)grid), dim3(256), lds, stream, a.x, L, \
                     s, D, logC, a.tw, a.epi)
    if (a.two_level) {
      if (epi) IMP_LAUNCH_DIT(true, true);
      else IMP_LAUNCH_DIT(true, false);
    } else {
      if (epi) IMP_LAUNCH_DIT(false, true);
      else IMP_LAUNCH_DIT(false, false);
    }
#undef IMP_LAUNCH_DIT
    IMP_CHECK(hipGetLastError());
    s += D;
  }
  return hipSuccess;
}

static bool plan_from(const unsigned int* passes, unsigned int n, u32 L, PassPlan* p) {
  if (!passes || n == 0) {
    default_plan(L, p);
    return true;
  }
  if (n > 32) return false;
  u32 sum = 0;
  for (u32 i = 0; i < n; i++) {
    if (passes[i] == 0 || passes[i] > 10) return false;
    p->d[i] = passes[i];
    sum += passes[i];
  }
  p->n = n;
  return sum == L;
}

// factor tables in stream-ordered temporary memory
static hipError_t make_factor(const u32* base_d, const u32* scale_d, u32 L, hipStream_t stream,
                              Factor* f, u32** tmp) {
  u32 ndig = L <= 10 ? 1 : (L <= 20 ? 2 : 3);
  IMP_CHECK(hipMallocAsync((void**)tmp, (size_t)3 * 1024 * 32, stream));
  hipLaunchKernelGGL(k_factor_tables, dim3((1024 * ndig + 255) / 256), dim3(256), 0, stream,
                     *tmp, base_d, scale_d, ndig);
  IMP_CHECK(hipGetLastError());
  f->tab = *tmp;
  f->ndig = ndig;
  return hipSuccess;
}

static std::mutex g_host_mu;

}  // namespace imp

using namespace imp;

// =============================================================== C ABI
extern "C" hipError_t ssip_ntt(PolyPtr x, const unsigned int* twiddle, unsigned int log_len,
                               hipStream_t stream, const unsigned int, const unsigned int) {
  IMP_REQUIRE(log_len <= 32, "log_len");
  IMP_REQUIRE(x.len == (1ull << log_len), "x.len must be 2^log_len");
  IMP_REQUIRE(x.ptr != nullptr && (log_len == 0 || twiddle != nullptr), "null pointer");
  NttArgs a = {};
  IMP_REQUIRE(make_view(x, &a.x), "PolyPtr view");
  a.L = log_len;
  a.tw.tab = twiddle;
  a.two_level = false;
  PassPlan plan;
  default_plan(log_len, &plan);
  a.plan = &plan;
  return run_ntt(a, stream);
}

extern "C" hipError_t recompute_ntt(PolyPtr x, const unsigned int* pq_d, unsigned int pq_deg,
                                    const unsigned int* omegas_d, unsigned int log_len,
                                    hipStream_t stream, const unsigned int, const unsigned int) {
  (void)pq_d;
  (void)pq_deg;
  IMP_REQUIRE(log_len <= 31, "log_len");
  IMP_REQUIRE(x.len == (1ull << log_len), "x.len must be 2^log_len");
  IMP_REQUIRE(x.ptr != nullptr && omegas_d != nullptr, "null pointer");
  NttArgs a = {};
  IMP_REQUIRE(make_view(x, &a.x), "PolyPtr view");
  a.L = log_len;
  PassPlan plan;
  default_plan(log_len, &plan);
  a.plan = &plan;
  u32* tabs = nullptr;
  if (log_len >= 2) {
    // two-level twiddles: lo[i] = w^i (i < 2^h), hi[j] = w^(j * 2^h)
    u32 h = (log_len - 1 + 1) / 2;
    u64 nlo = 1ull << h, nhi = 1ull << (log_len - 1 - h);
    IMP_CHECK(hipMallocAsync((void**)&tabs, (size_t)(nlo + nhi) * 32, stream));
    hipLaunchKernelGGL(k_pow_from_omegas, dim3((u32)((nlo + 255) / 256)), dim3(256), 0, stream,
                       tabs, nlo, (u64)0, omegas_d);
    hipLaunchKernelGGL(k_pow_from_omegas, dim3((u32)((nhi + 255) / 256)), dim3(256), 0, stream,
                       tabs + 8 * nlo, nhi, (u64)h, omegas_d);
    IMP_CHECK(hipGetLastError());
    a.two_level = true;
    a.tw.lo = tabs;
    a.tw.hi = tabs + 8 * nlo;
    a.tw.h = h;
  } else {
    a.tw.tab = omegas_d;  // n <= 2: only omega^0 = 1 is ever used
  }
  hipError_t e = run_ntt(a, stream);
  if (tabs) IMP_CHECK(hipFreeAsync(tabs, stream));
  return e;
}

extern "C" hipError_t ssip_precompute(unsigned int* twiddle, unsigned int len,
                                      const unsigned int* unit) {
  if (len == 0) return hipSuccess;
  IMP_REQUIRE(twiddle && unit, "null pointer");
  u32 *d_out = nullptr, *d_unit = nullptr;
  IMP_CHECK(hipMalloc((void**)&d_out, (size_t)len * 32));
  IMP_CHECK(hipMalloc((void**)&d_unit, 32));
  IMP_CHECK(hipMemcpy(d_unit, unit, 32, hipMemcpyHostToDevice));
  u32 chunk = 64;
  u64 threads = ((u64)len + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_powers, dim3((u32)((threads + 255) / 256)), dim3(256), 0, 0, d_out,
                     (u64)len, d_unit, chunk);
  IMP_CHECK(hipGetLastError());
  IMP_CHECK(hipMemcpy(twiddle, d_out, (size_t)len * 32, hipMemcpyDeviceToHost));
  IMP_CHECK(hipFree(d_out));
  IMP_CHECK(hipFree(d_unit));
  return hipSuccess;
}

extern "C" void gen_pq_omegas(unsigned int* pq, unsigned int* omegas, unsigned int pq_deg,
                              unsigned int len, unsigned int* unit) {
  // precompute.cuh:89-112 semantics (host outputs), computed on the device:
  // omegas[i] = unit^(2^i), i < 32; pq[i] = unit^(i * (len >> pq_deg)).
  if (!pq || !omegas || !unit || pq_deg == 0 || pq_deg > 31) return;
  if (((u64)len >> pq_deg) == 0) return;  // the reference requires len >= 2^pq_deg
  std::lock_guard<std::mutex> lk(g_host_mu);
  u64 npq = 1ull << (pq_deg - 1);
  u64 step = (u64)len >> pq_deg;
  u32 shift = 0;
  while ((1ull << shift) < step) shift++;
  u32* d = nullptr;
  if (hipMalloc((void**)&d, (size_t)(npq + 32) * 32) != hipSuccess) return;
  u32* d_om = d + 8 * npq;
  hipMemcpy(d_om, unit, 32, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_square_chain, dim3(1), dim3(1), 0, 0, d_om, 32u);
  hipLaunchKernelGGL(k_pow_from_omegas, dim3((u32)((npq + 255) / 256)), dim3(256), 0, 0, d, npq,
                     (u64)shift, d_om);
  hipMemcpy(omegas, d_om, 32 * 32, hipMemcpyDeviceToHost);
  hipMemcpy(pq, d, (size_t)npq * 32, hipMemcpyDeviceToHost);
  hipFree(d);
}

extern "C" hipError_t distribute_powers(PolyPtr poly, const unsigned int* powers,
                                        unsigned long long power_num, hipStream_t stream,
                                        const PolySlice* poly_slice) {
  IMP_REQUIRE(poly.ptr && powers && power_num > 0, "null pointer / power_num");
  View v;
  IMP_REQUIRE(make_view(poly, &v), "PolyPtr view");
  PolySlice sl;
  if (poly_slice) sl = *poly_slice;
  else sl = PolySlice{0, poly.len, poly.len};
  IMP_REQUIRE(sl.len > 0, "slice len");
  if (poly.len == 0) return hipSuccess;
  u64 grid = (poly.len + 255) / 256;
  hipLaunchKernelGGL(k_distribute_powers, dim3((u32)grid), dim3(256), 0, stream, v,
                     (u64)poly.len, powers, (u64)power_num, sl);
  return hipGetLastError();
}

extern "C" hipError_t impeller_twiddles(uint32_t* twiddle_d, unsigned long long len,
                                        const uint32_t* omega_host, hipStream_t stream) {
  if (len == 0) return hipSuccess;
  IMP_REQUIRE(twiddle_d && omega_host, "null pointer");
  u32* d_unit = nullptr;
  IMP_CHECK(hipMallocAsync((void**)&d_unit, 32, stream));
  IMP_CHECK(hipMemcpyAsync(d_unit, omega_host, 32, hipMemcpyHostToDevice, stream));
  u32 chunk = 64;
  u64 threads = (len + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_powers, dim3((u32)((threads + 255) / 256)), dim3(256), 0, stream,
                     twiddle_d, (u64)len, d_unit, chunk);
  IMP_CHECK(hipGetLastError());
  IMP_CHECK(hipFreeAsync(d_unit, stream));
  // the host buffer may go away when we return
  return hipStreamSynchronize(stream);
}

extern "C" hipError_t impeller_ntt(uint32_t* x, const uint32_t* twiddle, unsigned int log_len,
                                   int inverse, const uint32_t* coset_shift,
                                   const unsigned int* plan_passes, unsigned int n_passes,
                                   hipStream_t stream) {
  IMP_REQUIRE(log_len <= 31, "log_len");
  IMP_REQUIRE(x && (log_len == 0 || twiddle), "null pointer");
  NttArgs a = {};
  a.x = View{x, 0, 1ull << log_len};
  a.L = log_len;
  a.tw.tab = twiddle;
  PassPlan plan;
  IMP_REQUIRE(plan_from(plan_passes, n_passes, log_len, &plan), "pass plan");
  a.plan = &plan;
  u32* pro_tab = nullptr;
  u32* epi_tab = nullptr;
  u32* ninv = nullptr;
  if (!inverse) {
    if (coset_shift) IMP_CHECK(make_factor(coset_shift, nullptr, log_len, stream, &a.pro, &pro_tab));
  } else {
    IMP_CHECK(hipMallocAsync((void**)&ninv, 32, stream));
    hipLaunchKernelGGL(k_ninv, dim3(1), dim3(1), 0, stream, ninv, (u32)log_len);
    IMP_CHECK(hipGetLastError());
    // factor(i) = n^-1 * shift^i (shift given as its inverse by the caller)
    // or the constant n^-1 (base 1, one digit)
    IMP_CHECK(make_factor(coset_shift, ninv, coset_shift ? log_len : 10, stream, &a.epi,
                          &epi_tab));
  }
  hipError_t e = run_ntt(a, stream);
  if (pro_tab) IMP_CHECK(hipFreeAsync(pro_tab, stream));
  if (epi_tab) IMP_CHECK(hipFreeAsync(epi_tab, stream));
  if (ninv) IMP_CHECK(hipFreeAsync(ninv, stream));
  return e;
}

