// poly.hip -- the reference's polynomial library over BN254 Fr
// (core/src/poly/src/poly.h:6-30): mixed-length add / sub / mul, constant
// fills, scalar inverse and power, the scan-shaped operations scan_mul,
// batched_invert, poly_eval and kate_division, and the lookup permutation
// (permute) over a 256-bit-key radix sort.
//
// The scan-shaped operations share one two-pass tree.  A work-group takes a
// tile of STILE = SB * SK consecutive elements; thread t owns the run
// [t SK, t SK + SK).  The up pass writes each run's summary and the tile's
// summary (an element of the next level); levels repeat until one tile holds
// the level.  The down pass, from the top level back to level 0, combines a
// tile's carry with its run summaries (wave 0, through LDS) into each run's
// carry, from which every thread finishes its run.  Summaries:
//   product scans (scan_mul, batched_invert): the product of the run;
//   suffix Horner (kate_division, poly_eval): h = sum_j x_j b^j over the run,
//     combined as h_left + b^len(left) h_right.
// No work-group waits on another one; the levels are separate launches.
#include <algorithm>

#include "common.h"

namespace imp {

__device__ __forceinline__ Fr fmul(const Fr& a, const Fr& b) { return mul<FrParams>(a, b); }
__device__ __forceinline__ Fr fadd(const Fr& a, const Fr& b) { return add<FrParams>(a, b); }

__device__ __forceinline__ Fr vld(const View& v, u64 i) {
  return load_fe<Fr>(v.base + 8 * view_phys(v, i));
}
__device__ __forceinline__ void vst(const View& v, u64 i, const Fr& x) {
  store_fe<Fr>(v.base + 8 * view_phys(v, i), x);
}

static inline bool make_cview(const ConstPolyPtr& c, View* v) {
  PolyPtr p;
  p.ptr = const_cast<u32*>(c.ptr);
  p.len = c.len;
  p.rotate = c.rotate;
  p.offset = c.offset;
  p.whole_len = c.whole_len;
  return make_view(p, v);
}

static inline View flat_view(u32* p, u64 n) {
  View v;
  v.base = p;
  v.off = 0;
  v.wl = n ? n : 1;
  return v;
}

static hipError_t ew_grid(u64 len, u32* blocks) {
  int dev = 0, cus = 256;
  IMP_CHECK(hipGetDevice(&dev));
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  u64 b = (len + 255) / 256, cap = (u64)cus * 16;
  *blocks = (u32)(b == 0 ? 1 : (b > cap ? cap : b));
  return hipSuccess;
}

// ------------------------------------------------------------ element-wise
// poly_basic.cuh:4-56: r[i] = a[i] op b[i] for i < len, and for
// len <= i < upper the longer operand alone (negated when it is the
// subtrahend).  a is the longer operand (the entry points swap).
enum { EW_ADD = 0, EW_SUB = 1, EW_RSUB = 2, EW_MUL = 3 };

__global__ __launch_bounds__(256) void k_poly_binary(int op, View a, View b, View r, u64 len,
                                                     u64 upper) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < upper; i += stride) {
    const Fr x = vld(a, i);
    Fr z;
    if (i < len) {
      const Fr y = vld(b, i);
      switch (op) {
        case EW_ADD: z = fadd(x, y); break;
        case EW_SUB: z = sub<FrParams>(x, y); break;
        case EW_RSUB: z = sub<FrParams>(y, x); break;
        default: z = fmul(x, y);
      }
    } else {
      z = op == EW_RSUB ? neg<FrParams>(x) : x;
    }
    vst(r, i, z);
  }
}

// poly_basic.cuh:58-72: every element := value, or (Coef form) value where
// slice.to_origin(i) == 0 and zero elsewhere.  mode 0: zero, 1: one
// everywhere, 2: one at the origin.
__global__ __launch_bounds__(256) void k_poly_fill(int mode, View r, u64 len, u64 s_begin,
                                                   u64 s_len) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
    bool one = mode == 1;
    if (mode == 2) one = s_len == 0 ? i == 0 : (s_begin + i) % s_len == 0;
    vst(r, i, one ? Fr::one() : Fr::zero());
  }
}

// poly_basic.cuh:74-82 (one thread)
__global__ void k_inv_scalar(u32* t) { store_fe<Fr>(t, inv<FrParams>(load_fe<Fr>(t))); }

// field.cuh:1144-1159: t^e, square-and-multiply from the top bit
__global__ void k_scalar_pow(u32* t, u64 e) {
  const Fr a = load_fe<Fr>(t);
  Fr acc = Fr::one();
  for (int i = 63; i >= 0; i--) {
    acc = fmul(acc, acc);
    if ((e >> i) & 1) acc = fmul(acc, a);
  }
  store_fe<Fr>(t, acc);
}

// ------------------------------------------------------------ scan machinery
constexpr u32 SB = 256;          // threads per work-group
constexpr u32 SK = 8;            // elements per thread run
constexpr u32 STILE = SB * SK;   // elements per tile
constexpr u32 SLOG = 11;         // log2(STILE)
constexpr u32 SPW = SLOG + 1;    // powers per level: b^(2^e), e = 0..SLOG
constexpr u32 SMAXL = 6;         // levels (n < 2^64)

enum ScanOp { OP_PROD = 0, OP_INV = 1, OP_HORNER = 2 };

__device__ __forceinline__ Fr shfl_fe(const Fr& a, int src) {
  Fr r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.v[i] = (u32)__shfl((int)a.v[i], src, 64);
  return r;
}

__device__ __forceinline__ Fr sel(bool c, const Fr& a, const Fr& b) {
  Fr r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.v[i] = c ? a.v[i] : b.v[i];
  return r;
}

// thread run [first, first + SK) of a level; elements past n read as the
// operator's neutral padding (one for products, zero for the suffix Horner,
// where trailing zeros change no suffix value)
template <int OP>
__device__ __forceinline__ void load_run(const View& v, u64 n, u64 first, Fr x[SK]) {
#pragma unroll
  for (u32 j = 0; j < SK; j++) {
    const u64 e = first + j;
    x[j] = e < n ? vld(v, e) : (OP == OP_HORNER ? Fr::zero() : Fr::one());
  }
}

// batched_invert multiplies with zeros replaced by one (batched_invert.cuh:14-30)
__device__ __forceinline__ Fr nz(const Fr& x) { return x.is_zero() ? Fr::one() : x; }

template <int OP>
__device__ __forceinline__ Fr run_summary(const Fr x[SK], const Fr& beta) {
  Fr h;
  if (OP == OP_HORNER) {
    h = x[SK - 1];
#pragma unroll
    for (int j = SK - 2; j >= 0; j--) h = fadd(x[j], fmul(beta, h));
  } else {
    h = OP == OP_INV ? nz(x[0]) : x[0];
#pragma unroll
    for (u32 j = 1; j < SK; j++) h = fmul(h, OP == OP_INV ? nz(x[j]) : x[j]);
  }
  return h;
}

// up pass: run summaries -> lane_sum[tile * SB + t]; tile summary -> tile_sum[tile].
// pw: this level's powers b^(2^e) (suffix Horner only).
template <int OP>
__global__ __launch_bounds__(SB) void k_scan_up(View in, u64 n, u32* lane_sum, u32* tile_sum,
                                                const u32* pw) {
  __shared__ uint4 sh[SB * 2];
  u32* shw = reinterpret_cast<u32*>(sh);
  const u32 t = threadIdx.x;
  const u64 tile = blockIdx.x;
  Fr beta = Fr::zero();
  if (OP == OP_HORNER) beta = load_fe<Fr>(pw);
  Fr x[SK];
  load_run<OP>(in, n, tile * STILE + (u64)t * SK, x);
  const Fr h = run_summary<OP>(x, beta);
  store_fe<Fr>(lane_sum + 8 * (tile * SB + t), h);
  store_fe<Fr>(shw + 8 * t, h);
  __syncthreads();
  if (t >= 64) return;
  const u32 l = t;
  const Fr s0 = load_fe<Fr>(shw + 8 * (4 * l)), s1 = load_fe<Fr>(shw + 8 * (4 * l + 1));
  const Fr s2 = load_fe<Fr>(shw + 8 * (4 * l + 2)), s3 = load_fe<Fr>(shw + 8 * (4 * l + 3));
  Fr T;
  if (OP == OP_HORNER) {
    const Fr bk = load_fe<Fr>(pw + 8 * 3);  // b^SK
    T = fadd(s0, fmul(bk, fadd(s1, fmul(bk, fadd(s2, fmul(bk, s3))))));
    // lane l (a multiple of 2^(s+1)) accumulates sum_i T_(l+i) (b^(4 SK))^i
#pragma unroll
    for (u32 s = 0; s < 6; s++) {
      const Fr o = shfl_fe(T, (int)((l + (1u << s)) & 63));
      T = fadd(T, fmul(load_fe<Fr>(pw + 8 * (5 + s)), o));
    }
  } else {
    T = fmul(fmul(s0, s1), fmul(s2, s3));
#pragma unroll
    for (u32 s = 0; s < 6; s++) T = fmul(T, shfl_fe(T, (int)(l ^ (1u << s))));
  }
  if (l == 0) store_fe<Fr>(tile_sum + 8 * tile, T);
}

// down pass.  carry (carry2): this tile's carry from the level above, NULL at
// the top level, where the carry is the neutral one (products: one; suffix
// Horner: zero; batched_invert: prefix one and, for the inverse of the
// inclusive prefix, the inverse of the tile's own total).  LEVEL0 writes the
// operation's result (scan_mul: inclusive prefix products; batched_invert:
// x^-1, zero kept; kate: exclusive suffix Horner values) through out;
// otherwise out (out2) receive the carries of the level below: exclusive
// prefix products (and inverses of inclusive prefix products) or exclusive
// suffix Horner values.
template <int OP, bool LEVEL0>
__global__ __launch_bounds__(SB) void k_scan_down(View in, u64 n, const u32* lane_sum,
                                                  const u32* carry, const u32* carry2, View out,
                                                  View out2, const u32* pw) {
  __shared__ uint4 sh[SB * 2 * (OP == OP_INV ? 2 : 1)];
  u32* shw = reinterpret_cast<u32*>(sh);
  const u32 t = threadIdx.x;
  const u64 tile = blockIdx.x;
  const u64 first = tile * STILE + (u64)t * SK;
  Fr x[SK];
  load_run<OP>(in, n, first, x);
  store_fe<Fr>(shw + 8 * t, load_fe<Fr>(lane_sum + 8 * (tile * SB + t)));
  __syncthreads();
  if (t < 64) {
    const u32 l = t;
    const Fr s0 = load_fe<Fr>(shw + 8 * (4 * l)), s1 = load_fe<Fr>(shw + 8 * (4 * l + 1));
    const Fr s2 = load_fe<Fr>(shw + 8 * (4 * l + 2)), s3 = load_fe<Fr>(shw + 8 * (4 * l + 3));
    if (OP == OP_HORNER) {
      const Fr C = carry ? load_fe<Fr>(carry + 8 * tile) : Fr::zero();
      const Fr bk = load_fe<Fr>(pw + 8 * 3);
      const Fr b2 = fadd(s2, fmul(bk, s3));
      const Fr b1 = fadd(s1, fmul(bk, b2));
      Fr T = fadd(s0, fmul(bk, b1));
      // the carry enters after the last lane's span of 4 SK elements
      const Fr Tc = fadd(T, fmul(load_fe<Fr>(pw + 8 * 5), C));
      T = sel(l == 63, Tc, T);
#pragma unroll
      for (u32 s = 0; s < 6; s++) {
        const Fr o = shfl_fe(T, (int)((l + (1u << s)) & 63));
        const Fr Ts = fadd(T, fmul(load_fe<Fr>(pw + 8 * (5 + s)), o));
        T = sel(l + (1u << s) < 64, Ts, T);
      }
      Fr X = shfl_fe(T, (int)((l + 1) & 63));
      X = sel(l == 63, C, X);
      const Fr e2 = fadd(s3, fmul(bk, X));
      const Fr e1 = fadd(s2, fmul(bk, e2));
      const Fr e0 = fadd(s1, fmul(bk, e1));
      store_fe<Fr>(shw + 8 * (4 * l), e0);
      store_fe<Fr>(shw + 8 * (4 * l + 1), e1);
      store_fe<Fr>(shw + 8 * (4 * l + 2), e2);
      store_fe<Fr>(shw + 8 * (4 * l + 3), X);
    } else {
      // forward exclusive products with the carry in front
      const Fr C = carry ? load_fe<Fr>(carry + 8 * tile) : Fr::one();
      const Fr a1 = fmul(s0, s1);
      const Fr a2 = fmul(a1, s2);
      Fr T = fmul(a2, s3);
      T = sel(l == 0, fmul(C, T), T);
#pragma unroll
      for (u32 s = 0; s < 6; s++) {
        const Fr o = shfl_fe(T, (int)((l - (1u << s)) & 63));
        const Fr Ts = fmul(o, T);
        T = sel(l >= (1u << s), Ts, T);
      }
      Fr X = shfl_fe(T, (int)((l - 1) & 63));
      X = sel(l == 0, C, X);
      const Fr p1 = fmul(X, s0), p2 = fmul(X, a1), p3 = fmul(X, a2);
      if (OP == OP_INV) {
        // suffix exclusive products, times the inverse of the inclusive
        // prefix at the tile's end
        Fr IC;
        if (carry2) {
          IC = load_fe<Fr>(carry2 + 8 * tile);
        } else {
          IC = inv<FrParams>(shfl_fe(T, 63));  // top level: the total's inverse
        }
        const Fr c1 = fmul(s2, s3);
        const Fr c0 = fmul(s1, c1);
        Fr U = fmul(s0, c0);
#pragma unroll
        for (u32 s = 0; s < 6; s++) {
          const Fr o = shfl_fe(U, (int)((l + (1u << s)) & 63));
          const Fr Us = fmul(U, o);
          U = sel(l + (1u << s) < 64, Us, U);
        }
        Fr Y = shfl_fe(U, (int)((l + 1) & 63));
        Y = sel(l == 63, IC, fmul(Y, IC));
        const Fr q2 = fmul(Y, s3), q1 = fmul(Y, c1), q0 = fmul(Y, c0);
        u32* sh2 = shw + 8 * SB;
        store_fe<Fr>(sh2 + 8 * (4 * l), q0);
        store_fe<Fr>(sh2 + 8 * (4 * l + 1), q1);
        store_fe<Fr>(sh2 + 8 * (4 * l + 2), q2);
        store_fe<Fr>(sh2 + 8 * (4 * l + 3), Y);
      }
      store_fe<Fr>(shw + 8 * (4 * l), X);
      store_fe<Fr>(shw + 8 * (4 * l + 1), p1);
      store_fe<Fr>(shw + 8 * (4 * l + 2), p2);
      store_fe<Fr>(shw + 8 * (4 * l + 3), p3);
    }
  }
  __syncthreads();
  const Fr E = load_fe<Fr>(shw + 8 * t);
  if (OP == OP_HORNER) {
    const Fr beta = load_fe<Fr>(pw);
    Fr Q = E;
#pragma unroll
    for (int j = SK - 1; j >= 0; j--) {
      if (first + j < n) vst(out, first + j, Q);
      if (j > 0) Q = fadd(x[j], fmul(beta, Q));
    }
  } else if (OP == OP_PROD) {
    Fr acc = E;
#pragma unroll
    for (u32 j = 0; j < SK; j++) {
      if (LEVEL0) {
        acc = fmul(acc, x[j]);
        if (first + j < n) vst(out, first + j, acc);
      } else {
        if (first + j < n) vst(out, first + j, acc);
        if (j + 1 < SK) acc = fmul(acc, x[j]);
      }
    }
  } else {
    // p_j = prefix before j, i_j = inverse of the prefix through j
    Fr p[SK];
    p[0] = E;
#pragma unroll
    for (u32 j = 1; j < SK; j++) p[j] = fmul(p[j - 1], nz(x[j - 1]));
    Fr I = load_fe<Fr>(shw + 8 * SB + 8 * t);
#pragma unroll
    for (int j = SK - 1; j >= 0; j--) {
      if (first + j < n) {
        if (LEVEL0) {
          vst(out, first + j, x[j].is_zero() ? Fr::zero() : fmul(p[j], I));
        } else {
          vst(out, first + j, p[j]);
          vst(out2, first + j, I);
        }
      }
      if (j > 0) I = fmul(I, nz(x[j]));
    }
  }
}

// pw[SPW l + e] = b^(2^(SLOG l + e)) for l < levels, e <= SLOG (one lane)
__global__ void k_pow2_chain(const u32* b, u32* pw, u32 levels) {
  Fr v = load_fe<Fr>(b);
  for (u32 l = 0; l < levels; l++) {
    for (u32 e = 0; e < SLOG; e++) {
      store_fe<Fr>(pw + 8 * (SPW * l + e), v);
      v = fmul(v, v);
    }
    store_fe<Fr>(pw + 8 * (SPW * l + SLOG), v);
  }
}

__global__ void k_store_zero(u32* r) { store_fe<Fr>(r, Fr::zero()); }

// ------------------------------------------------ u32 exclusive scan
// Device-wide exclusive sums of u32 counts, tiles of STILE values (8 per
// thread), with the same up / down level structure as the Fr scans: the up
// pass writes tile sums (the next level's values), the down pass scans a
// tile in place from its carry (the scanned tile sum of the level above).
__device__ __forceinline__ u32 block_excl_u32(u32 v, u32* sh, u32* total) {
  const u32 t = threadIdx.x, lane = t & 63, w = t >> 6;
  u32 inc = v;
#pragma unroll
  for (u32 s = 1; s < 64; s <<= 1) {
    const u32 o = (u32)__shfl((int)inc, (int)((lane - s) & 63), 64);
    if (lane >= s) inc += o;
  }
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  u32 before = 0, all = 0;
#pragma unroll
  for (u32 k = 0; k < SB / 64; k++) {
    const u32 x = sh[k];
    if (k < w) before += x;
    all += x;
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

// guard (optional): the launch does nothing when *guard == 0 (a radix pass
// the sort's device-side schedule skips)
__global__ __launch_bounds__(SB) void k_u32_up(const u32* in, u64 n, u32* tile_sum,
                                               const u32* guard) {
  if (guard && *guard == 0) return;
  __shared__ u32 sh[SB / 64];
  const u32 t = threadIdx.x;
  const u64 first = (u64)blockIdx.x * STILE + (u64)t * SK;
  u32 s = 0;
#pragma unroll
  for (u32 j = 0; j < SK; j++)
    if (first + j < n) s += in[first + j];
  u32 total;
  block_excl_u32(s, sh, &total);
  if (t == 0) tile_sum[blockIdx.x] = total;
}

__global__ __launch_bounds__(SB) void k_u32_down(u32* data, u64 n, const u32* carry,
                                                 const u32* guard) {
  if (guard && *guard == 0) return;
  __shared__ u32 sh[SB / 64];
  const u32 t = threadIdx.x;
  const u64 first = (u64)blockIdx.x * STILE + (u64)t * SK;
  u32 x[SK], s = 0;
#pragma unroll
  for (u32 j = 0; j < SK; j++) {
    x[j] = first + j < n ? data[first + j] : 0u;
    s += x[j];
  }
  u32 total;
  u32 acc = block_excl_u32(s, sh, &total) + (carry ? carry[blockIdx.x] : 0u);
#pragma unroll
  for (u32 j = 0; j < SK; j++) {
    if (first + j < n) data[first + j] = acc;
    acc += x[j];
  }
}

static size_t u32scan_bytes(u64 n) {
  size_t b = 0;
  for (u64 m = n; m > STILE;) {
    m = (m + STILE - 1) / STILE;
    b += align_up((size_t)m * 4, 256);
  }
  return b;
}

static hipError_t u32scan(u32* data, u64 n, void* temp, hipStream_t stream,
                          const u32* guard = nullptr) {
  if (n == 0) return hipSuccess;
  u32* lv[SMAXL + 1];
  u64 ln[SMAXL + 1];
  u32 L = 0;
  lv[0] = data;
  ln[0] = n;
  char* T = static_cast<char*>(temp);
  while (ln[L] > STILE) {
    const u64 tiles = (ln[L] + STILE - 1) / STILE;
    lv[L + 1] = reinterpret_cast<u32*>(T);
    T += align_up((size_t)tiles * 4, 256);
    ln[L + 1] = tiles;
    hipLaunchKernelGGL(k_u32_up, dim3((u32)tiles), dim3(SB), 0, stream, lv[L], ln[L],
                       lv[L + 1], guard);
    IMP_CHECK(hipGetLastError());
    L++;
  }
  for (int l = (int)L; l >= 0; l--) {
    const u64 tiles = (ln[l] + STILE - 1) / STILE;
    hipLaunchKernelGGL(k_u32_down, dim3((u32)tiles), dim3(SB), 0, stream, lv[l], ln[l],
                       l == (int)L ? (const u32*)nullptr : lv[l + 1], guard);
    IMP_CHECK(hipGetLastError());
  }
  return hipSuccess;
}

// ------------------------------------------------- 256-bit key radix sort
// LSD over the 32 bytes of canonical 256-bit numbers (32-byte records),
// ascending, stable.  Per pass: a digit histogram per tile (digit-major, so
// one exclusive scan gives every tile's stable output offsets), then a
// stable scatter (k_rs_scatter_split: the tile ordered by the digit in LDS,
// written out by digit runs).
//
// Fewer passes and fewer bytes where the keys allow, decided on the device
// (no host synchronisation).  Prefix phase (slots 0..7, passes 24..31): the
// keys' top 8 bytes and their input index, 12-byte records, are sorted, and
// the keys are then gathered in that order.  When no two different keys
// share those bytes -- the rule for values spread over the field -- that
// order is already the full order (equal keys keep their input order, as
// the full sort keeps it), and the 32 slots of the full phase (passes 0..31
// over the gathered keys) do nothing.  Otherwise they run, an LSD sort of
// the whole key.  In both phases a byte equal in every key (the AND and the
// OR of all keys agree there, e.g. the high bytes of small integers) needs
// no pass.  A slot's kernels read the schedule word (run, source buffer) the
// plan kernel wrote and return at once for a skipped pass; a last kernel
// moves the result to its buffer when an odd number of full passes ran.
constexpr u32 RPASSES = 32;
constexpr u32 RPREFIX = 8;                   // slots of the prefix phase (bytes 24..31)
constexpr u32 RSLOTS = RPREFIX + RPASSES;
constexpr u32 RS_PAR_PREFIX = 2 * RSLOTS;    // record buffer (0 or 1) after the prefix phase
constexpr u32 RS_TIES = RS_PAR_PREFIX + 1;   // two different keys share the top 8 bytes
constexpr u32 RS_PAR_END = RS_TIES + 1;      // key buffer (0: result, 1: scratch) at the end
constexpr u32 RS_AND = RS_PAR_END + 1;       // 8 words: AND over all keys
constexpr u32 RS_OR = RS_AND + 8;            // 8 words: OR over all keys
constexpr u32 RS_WORDS = RS_OR + 8;
constexpr u32 RS_GRID = 1024;                // work-groups of the tile loops
constexpr u32 RREC = 3;                      // words per prefix record: key words 6, 7, index
static_assert(RS_WORDS <= 128, "k_rs_init runs 128 threads, one per schedule word");

__host__ __device__ inline u32 slot_pass(u32 slot) {
  return slot < RPREFIX ? 32 - RPREFIX + slot : slot - RPREFIX;
}

__global__ void k_rs_init(u32* sched) {
  const u32 t = threadIdx.x;
  if (t < RS_WORDS) sched[t] = (t >= RS_AND && t < RS_OR) ? 0xffffffffu : 0u;
}

// AND and OR of all keys, word by word, and each key's prefix record
__global__ __launch_bounds__(SB) void k_rs_bits(const u32* keys, u64 n, u32* sched, u32* rec) {
  __shared__ u32 sa[8], so[8];
  const u32 t = threadIdx.x;
  if (t < 8) {
    sa[t] = 0xffffffffu;
    so[t] = 0u;
  }
  u32 a[8], o[8];
#pragma unroll
  for (u32 w = 0; w < 8; w++) {
    a[w] = 0xffffffffu;
    o[w] = 0u;
  }
  for (u64 i = (u64)blockIdx.x * SB + t; i < n; i += (u64)gridDim.x * SB) {
    const Fr k = load_fe<Fr>(keys + 8 * i);
#pragma unroll
    for (u32 w = 0; w < 8; w++) {
      a[w] &= k.v[w];
      o[w] |= k.v[w];
    }
    rec[RREC * i] = k.v[6];
    rec[RREC * i + 1] = k.v[7];
    rec[RREC * i + 2] = (u32)i;
  }
  __syncthreads();
#pragma unroll
  for (u32 w = 0; w < 8; w++) {
    atomicAnd(&sa[w], a[w]);
    atomicOr(&so[w], o[w]);
  }
  __syncthreads();
  if (t < 8) {
    atomicAnd(&sched[RS_AND + t], sa[t]);
    atomicOr(&sched[RS_OR + t], so[t]);
  }
}

// the schedule of one phase (0: prefix slots, 1: the full slots, which start
// from the gathered keys in buffer 0)
__global__ void k_rs_plan(u32* sched, u32 phase) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  u32 par = 0;
  const bool need = phase == 0 || sched[RS_TIES] != 0;
  const u32 s0 = phase ? RPREFIX : 0u, s1 = phase ? RSLOTS : RPREFIX;
  for (u32 s = s0; s < s1; s++) {
    const u32 pass = slot_pass(s), w = pass >> 2, sh = (pass & 3) * 8;
    const bool varies = (((sched[RS_AND + w] ^ sched[RS_OR + w]) >> sh) & 255u) != 0;
    const u32 run = need && varies ? 1u : 0u;
    sched[2 * s] = run;
    sched[2 * s + 1] = par;
    par ^= run;
  }
  sched[phase ? RS_PAR_END : RS_PAR_PREFIX] = par;
}

// out[i] = keys[index of prefix record i]
__global__ __launch_bounds__(SB) void k_rs_gather(const u32* r0, const u32* r1, const u32* keys,
                                                  u64 n, const u32* sched, u32* out) {
  const u32* rec = sched[RS_PAR_PREFIX] ? r1 : r0;
  for (u64 i = (u64)blockIdx.x * SB + threadIdx.x; i < n; i += (u64)gridDim.x * SB)
    store_fe<Fr>(out + 8 * i, load_fe<Fr>(keys + 8 * (u64)rec[RREC * i + 2]));
}

// RS_TIES when two adjacent prefix-sorted keys agree in their top 8 bytes
// but not in the rest (one atomic per wave that found such a pair: on a
// table of small integers every pair does)
__global__ __launch_bounds__(SB) void k_rs_ties(const u32* k, u64 n, u32* sched) {
  bool tie = false;
  for (u64 i = 1 + (u64)blockIdx.x * SB + threadIdx.x; i < n; i += (u64)gridDim.x * SB) {
    const u32* p = k + 8 * i;
    const u32* q = p - 8;
    if (p[7] == q[7] && p[6] == q[6]) {
#pragma unroll
      for (u32 w = 0; w < 6; w++) tie |= p[w] != q[w];
    }
  }
  const unsigned long long any = __ballot(tie);
  if (any != 0 && (threadIdx.x & 63) == (u32)(__ffsll((long long)any) - 1))
    atomicOr(&sched[RS_TIES], 1u);
}

// the pass's digit: byte (pass & 3) of key word pass >> 2 (prefix records
// hold key words 6 and 7 as their words 0 and 1)
template <u32 W>
__device__ __forceinline__ u32 rec_digit(u32 word, u32 pass) {
  return (word >> ((pass & 3) * 8)) & 255u;
}
template <u32 W>
__device__ __forceinline__ u32 digit_word(u32 pass) {
  return W == 8 ? (pass >> 2) : (pass >> 2) - 6;
}

template <u32 W>
__global__ __launch_bounds__(SB) void k_rs_hist(const u32* a, const u32* b, u64 n, u32 slot,
                                                const u32* sched, u32* cnt, u32 tiles) {
  if (sched[2 * slot] == 0) return;
  const u32* recs = sched[2 * slot + 1] ? b : a;
  const u32 pass = slot_pass(slot), dw = digit_word<W>(pass);
  __shared__ u32 h[256];
  const u32 t = threadIdx.x;
  for (u32 tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    h[t] = 0;
    __syncthreads();
#pragma unroll
    for (u32 j = 0; j < SK; j++) {
      const u64 i = (u64)tile * STILE + (u64)j * SB + t;
      if (i < n) atomicAdd(&h[rec_digit<W>(recs[W * i + dw], pass)], 1u);
    }
    __syncthreads();
    cnt[(u64)t * tiles + tile] = h[t];
    __syncthreads();
  }
}

// Prefix-record passes: each thread takes the run of SK = 8 consecutive
// records [8t, 8t + 8) of the tile; two stable 4-bit splits in LDS (counts
// per (field value, thread), one block scan, ranks) order the tile by the
// pass's digit, input order kept within a digit; then consecutive threads
// write consecutive positions of each digit's run, record i of digit d to
// gbase[d] + (i - dbase[d]) (gbase: the digit's scanned global start for
// this tile, dbase: its first staged position).
template <u32 W>
struct RsSplitLds {
  union {
    u32 cnt[16 * SB];  // split counts, then running ranks (swizzled)
    struct {
      u32 gbase[256];
      u32 dbase[256];
    } run;
  };
  u32 st[W][STILE];  // the staged tile, one plane per record word (swizzled)
};  // 80 KiB for 32-byte keys: two work-groups share a CU's 160 KiB

// LDS swizzles: a thread's run of 16 counters / 8 staged words lands in
// distinct banks across a wave, and 64 consecutive words stay distinct
__device__ __forceinline__ u32 cswz(u32 i) { return i ^ ((i >> 6) & 15u); }
__device__ __forceinline__ u32 sswz(u32 i) { return i ^ ((i >> 6) & 7u); }

template <u32 W>
__device__ __forceinline__ void rs_split(RsSplitLds<W>& L, const u32 (&r)[W][SK], u32 n, u32 dw,
                                         u32 sh) {
  const u32 t = threadIdx.x;
  // the digit words, picked with constant indices so r stays in registers
  u32 dg[SK];
  u32 sel[W];
#pragma unroll
  for (u32 w = 0; w < W; w++) sel[w] = 0u - (u32)(dw == w);
#pragma unroll
  for (u32 j = 0; j < SK; j++) {
    u32 x = 0;
#pragma unroll
    for (u32 w = 0; w < W; w++) x |= r[w][j] & sel[w];
    dg[j] = x;
  }
#pragma unroll
  for (u32 d = 0; d < 16; d++) L.cnt[cswz(d * SB + t)] = 0;
#pragma unroll
  for (u32 j = 0; j < SK; j++)
    if (j < n) L.cnt[cswz(((dg[j] >> sh) & 15u) * SB + t)] += 1;
  __syncthreads();
  u32 loc[16], sum = 0;
#pragma unroll
  for (u32 k = 0; k < 16; k++) {
    loc[k] = L.cnt[cswz(16 * t + k)];
    sum += loc[k];
  }
  u32 total;
  // the scan's four wave sums go to the first words of the staging planes:
  // no thread reads or writes those planes until the scan's last barrier
  // (each run was taken into registers before the split began)
  u32 run = block_excl_u32(sum, &L.st[0][0], &total);
#pragma unroll
  for (u32 k = 0; k < 16; k++) {
    L.cnt[cswz(16 * t + k)] = run;
    run += loc[k];
  }
  __syncthreads();
#pragma unroll
  for (u32 j = 0; j < SK; j++) {
    if (j < n) {
      const u32 c = cswz(((dg[j] >> sh) & 15u) * SB + t);
      const u32 rk = L.cnt[c];
      L.cnt[c] = rk + 1;
#pragma unroll
      for (u32 w = 0; w < W; w++) L.st[w][sswz(rk)] = r[w][j];
    }
  }
  __syncthreads();
}

template <u32 W>
__global__ __launch_bounds__(SB) void k_rs_scatter_split(u32* a, u32* b, u64 n, u32 slot,
                                                         const u32* sched, const u32* off,
                                                         u32 tiles) {
  if (sched[2 * slot] == 0) return;
  const u32* rin = sched[2 * slot + 1] ? b : a;
  u32* rout = sched[2 * slot + 1] ? a : b;
  const u32 pass = slot_pass(slot), dw = digit_word<W>(pass), sh = (pass & 3) * 8;
  __shared__ RsSplitLds<W> L;
  const u32 t = threadIdx.x;
  for (u32 tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const u64 tile0 = (u64)tile * STILE;
    const u32 n_tile = (u32)(n - tile0 < STILE ? n - tile0 : STILE);
    const u32 nr = n_tile > SK * t ? (n_tile - SK * t < SK ? n_tile - SK * t : SK) : 0;
    u32 r[W][SK];
    const u32* src = rin + W * (tile0 + (u64)SK * t);
    if (nr == SK) {
      // the run's W * SK words are 16-byte aligned (W * SK * 4 is a multiple of 16)
      const uint4* q = reinterpret_cast<const uint4*>(src);
#pragma unroll
      for (u32 k = 0; k < W * SK / 4; k++) {
        const uint4 x = q[k];
        const u32 e[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (u32 m = 0; m < 4; m++) r[(4 * k + m) % W][(4 * k + m) / W] = e[m];
      }
    } else {
#pragma unroll
      for (u32 j = 0; j < SK; j++)
#pragma unroll
        for (u32 w = 0; w < W; w++) r[w][j] = j < nr ? src[W * j + w] : 0u;
    }
    __syncthreads();  // the previous tile's writes out of LDS are done
    rs_split<W>(L, r, nr, dw, sh);
#pragma unroll
    for (u32 j = 0; j < SK; j++)
#pragma unroll
      for (u32 w = 0; w < W; w++) r[w][j] = L.st[w][sswz(SK * t + j)];
    __syncthreads();
    rs_split<W>(L, r, nr, dw, sh + 4);
    // digit run starts (the split counters are dead: gbase / dbase reuse them)
    L.run.gbase[t] = off[(u64)t * tiles + tile];
#pragma unroll
    for (u32 j = 0; j < SK; j++) {
      const u32 i = SK * t + j;
      if (i < n_tile) {
        const u32 d = (L.st[dw][sswz(i)] >> sh) & 255u;
        if (i == 0 || ((L.st[dw][sswz(i - 1)] >> sh) & 255u) != d) L.run.dbase[d] = i;
      }
    }
    __syncthreads();
    for (u32 i = t; i < n_tile; i += SB) {
      const u32 d = (L.st[dw][sswz(i)] >> sh) & 255u;
      const u64 pos = (u64)L.run.gbase[d] + i - L.run.dbase[d];
      if constexpr (W == 8) {
        Fr v;
#pragma unroll
        for (u32 w = 0; w < 8; w++) v.v[w] = L.st[w][sswz(i)];
        store_fe<Fr>(rout + 8 * pos, v);
      } else {
#pragma unroll
        for (u32 w = 0; w < W; w++) rout[W * pos + w] = L.st[w][sswz(i)];
      }
    }
  }
}

// the sorted keys into a when the last full pass that ran wrote b
__global__ __launch_bounds__(SB) void k_rs_finish(u32* a, const u32* b, u64 n, const u32* sched) {
  if (sched[RS_PAR_END] == 0) return;
  const uint4* src = reinterpret_cast<const uint4*>(b);
  uint4* dst = reinterpret_cast<uint4*>(a);
  for (u64 i = (u64)blockIdx.x * SB + threadIdx.x; i < 2 * n; i += (u64)gridDim.x * SB)
    dst[i] = src[i];
}

struct SortPlan {
  u32 tiles;
  size_t cnt, sched, scan, rec0, rec1, bytes;  // offsets within the sort's own area
};

static void sort_plan(u64 n, SortPlan* p) {
  p->tiles = (u32)((n + STILE - 1) / STILE);
  p->cnt = 0;
  p->sched = align_up((size_t)p->tiles * 256 * 4, 256);
  p->scan = p->sched + align_up(RS_WORDS * 4, 256);
  p->rec0 = p->scan + align_up(u32scan_bytes((u64)p->tiles * 256), 256);
  p->rec1 = p->rec0 + align_up((size_t)n * RREC * 4, 256);
  p->bytes = p->rec1 + align_up((size_t)n * RREC * 4, 256);
}

// sorts the n >= 1 keys in b into a (b is clobbered)
static hipError_t sort256(u32* a, u32* b, u64 n, void* area, hipStream_t stream) {
  SortPlan p;
  sort_plan(n, &p);
  char* A = static_cast<char*>(area);
  u32* cnt = reinterpret_cast<u32*>(A + p.cnt);
  u32* sched = reinterpret_cast<u32*>(A + p.sched);
  u32* r0 = reinterpret_cast<u32*>(A + p.rec0);
  u32* r1 = reinterpret_cast<u32*>(A + p.rec1);
  // The prefix passes (which run for all but constant bytes) launch one
  // work-group per tile, up to 8 x RS_GRID, so that many tiles' loads are in
  // flight at once (record histogram 79 -> 58 us, scatter 132 -> 123 us per
  // pass at 2^24); the full passes, which mostly return at once, stride
  // over the tiles with RS_GRID work-groups, so a skipped pass costs two
  // small launches.
  const u32 grid = std::min(p.tiles, RS_GRID);
  const u32 pgrid = std::min(p.tiles, 8 * RS_GRID);
  const u32 egrid = (u32)std::min<u64>((n + SB - 1) / SB, RS_GRID);
  auto pass = [&](u32 slot, auto hist, auto scatter, u32* x, u32* y) -> hipError_t {
    const u32 g = slot < RPREFIX ? pgrid : grid;
    hipLaunchKernelGGL(hist, dim3(g), dim3(SB), 0, stream, x, y, n, slot, sched, cnt, p.tiles);
    IMP_CHECK(hipGetLastError());
    IMP_CHECK(u32scan(cnt, (u64)p.tiles * 256, A + p.scan, stream, sched + 2 * slot));
    hipLaunchKernelGGL(scatter, dim3(g), dim3(SB), 0, stream, x, y, n, slot, sched, cnt,
                       p.tiles);
    return hipGetLastError();
  };
  hipLaunchKernelGGL(k_rs_init, dim3(1), dim3(128), 0, stream, sched);
  IMP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_rs_bits, dim3(egrid), dim3(SB), 0, stream, b, n, sched, r0);
  IMP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_rs_plan, dim3(1), dim3(64), 0, stream, sched, 0u);
  IMP_CHECK(hipGetLastError());
  for (u32 slot = 0; slot < RPREFIX; slot++)
    IMP_CHECK(pass(slot, k_rs_hist<RREC>, k_rs_scatter_split<RREC>, r0, r1));
  hipLaunchKernelGGL(k_rs_gather, dim3(egrid), dim3(SB), 0, stream, r0, r1, b, n, sched, a);
  IMP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_rs_ties, dim3(egrid), dim3(SB), 0, stream, a, n, sched);
  IMP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_rs_plan, dim3(1), dim3(64), 0, stream, sched, 1u);
  IMP_CHECK(hipGetLastError());
  for (u32 slot = RPREFIX; slot < RSLOTS; slot++)
    IMP_CHECK(pass(slot, k_rs_hist<8>, k_rs_scatter_split<8>, a, b));
  hipLaunchKernelGGL(k_rs_finish, dim3(egrid), dim3(SB), 0, stream, a, b, n, sched);
  return hipGetLastError();
}

// ---------------------------------------------------------- permute
// permute.cuh semantics: res_input = the usable inputs sorted by canonical
// value; res_table holds at each run start of res_input that value (one
// copy taken from the table) and at the other rows the table's leftover
// values in ascending order.  An input value missing from the table takes
// its own value at its run start (the reference asserts instead).
__global__ void k_to_number(View in, u64 n, u32* out) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) store_fe<Fr>(out + 8 * i, to_number<FrParams>(vld(in, i)));
}

__global__ void k_fill_u32(u32* p, u64 n, u32 v) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ int cmp_num(const Fr& a, const Fr& b) {
#pragma unroll
  for (int k = 7; k >= 0; k--) {
    if (a.v[k] != b.v[k]) return a.v[k] < b.v[k] ? -1 : 1;
  }
  return 0;
}

// run starts of the sorted inputs take their value and mark one table copy used
// first position of the sorted table st whose value is not below key
// (upper: above key) within [lo, hi)
__device__ __forceinline__ u64 table_bound(const u32* st, u64 lo, u64 hi, const Fr& key, bool upper) {
  while (lo < hi) {
    const u64 mid = lo + (hi - lo) / 2;
    const int c = cmp_num(load_fe<Fr>(st + 8 * mid), key);
    if (c < 0 || (upper && c == 0)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The table window of each block of 256 sorted inputs: win[2b] = first
// table position not below input 256 b, win[2b + 1] = first position above
// the block's last input (two threads per block, one search each).  Every
// input of the block finds its value's first table copy inside the window.
__global__ void k_perm_window(const u32* sa, const u32* st, u64 n, u32* win) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x, b = t >> 1;
  if (b * 256 >= n) return;
  const bool upper = (t & 1) != 0;
  const u64 k = upper ? (b * 256 + 256 < n ? b * 256 + 255 : n - 1) : b * 256;
  win[t] = (u32)table_bound(st, 0, n, load_fe<Fr>(sa + 8 * k), upper);
}

// run starts of the sorted inputs take their value and mark their value's
// first table copy used (the block's window narrows the search)
__global__ void k_perm_lookup(const u32* sa, const u32* st, u64 n, const u32* win, u32* start,
                              u32* avail, View res_table) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Fr a = load_fe<Fr>(sa + 8 * i);
  const bool s = i == 0 || cmp_num(a, load_fe<Fr>(sa + 8 * (i - 1))) != 0;
  start[i] = s ? 1u : 0u;
  if (!s) return;
  const u64 wl = win[2 * (i >> 8)], wh = win[2 * (i >> 8) + 1];
  const u64 p = table_bound(st, wl, wh, a, false);
  if (p < wh && cmp_num(load_fe<Fr>(st + 8 * p), a) == 0) avail[p] = 0;
  vst(res_table, i, from_number<FrParams>(a));
}

__global__ void k_perm_compact(const u32* st, const u32* avail, const u32* pos, u64 n,
                               u32* left) {
  const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n && avail[j]) store_fe<Fr>(left + 8 * (u64)pos[j], load_fe<Fr>(st + 8 * j));
}

__global__ void k_perm_fill(const u32* sa, const u32* left, const u32* nstart, u64 n,
                            View res_input, View res_table) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Fr a = load_fe<Fr>(sa + 8 * i);
  vst(res_input, i, from_number<FrParams>(a));
  const bool s = i == 0 || cmp_num(a, load_fe<Fr>(sa + 8 * (i - 1))) != 0;
  // nstart[i]: run starts before row i (exclusive sum), so the row is the
  // (i - nstart[i])-th repeated row
  if (!s) vst(res_table, i, from_number<FrParams>(load_fe<Fr>(left + 8 * (i - nstart[i]))));
}

struct PermPlan {
  size_t k0, k1, k2, start, avail, pos, sort, scan, bytes;
};

static void perm_plan(u64 n, PermPlan* p) {
  size_t off = 0;
  const size_t keys = align_up((size_t)n * 32, 256), flags = align_up((size_t)n * 4, 256);
  p->k0 = off;
  off += keys;
  p->k1 = off;
  off += keys;
  p->k2 = off;
  off += keys;
  p->start = off;
  off += flags;
  p->avail = off;
  off += flags;
  p->pos = off;
  off += flags;
  SortPlan sp;
  sort_plan(n, &sp);
  p->sort = off;
  p->scan = off;  // the flag scans reuse the sort's area after both sorts
  off += std::max(sp.bytes, u32scan_bytes(n));
  p->bytes = off;
}

// Temporary-buffer layout of a scan over n elements (n >= 1)
struct ScanPlan {
  u32 levels;
  u64 n[SMAXL];
  u64 tiles[SMAXL];
  size_t lane[SMAXL], sum[SMAXL], carry[SMAXL], carry2[SMAXL], pw;
  size_t bytes;
};

static void scan_plan(u64 n, int op, ScanPlan* s) {
  size_t off = 0;
  u32 l = 0;
  u64 m = n;
  for (;;) {
    const u64 t = (m + STILE - 1) / STILE;
    s->n[l] = m;
    s->tiles[l] = t;
    s->lane[l] = off;
    off += (size_t)t * SB * 32;
    s->sum[l] = off;
    off += (size_t)t * 32;
    s->carry[l] = s->carry2[l] = 0;
    if (t > 1) {
      s->carry[l] = off;
      off += (size_t)t * 32;
      if (op == OP_INV) {
        s->carry2[l] = off;
        off += (size_t)t * 32;
      }
    }
    l++;
    if (t == 1) break;
    m = t;
  }
  s->levels = l;
  s->pw = off;
  if (op == OP_HORNER) off += (size_t)SPW * l * 32;
  s->bytes = off;
}

static size_t scan_bytes(u64 n, int op) {
  if (n == 0) return 0;
  ScanPlan s;
  scan_plan(n, op, &s);
  return s.bytes;
}

// Runs the up passes and (down = true) the down passes of a scan of the
// view `in` (n elements) in the caller's temporary buffer.  base: the
// Horner base (device, suffix Horner only); top_sum (nullable) receives the
// whole input's summary.
template <int OP>
static hipError_t run_scan(const View& in, u64 n, const View& out, void* temp, size_t temp_bytes,
                           const u32* base, bool down, u32* top_sum, hipStream_t stream) {
  ScanPlan s;
  scan_plan(n, OP, &s);
  IMP_REQUIRE(temp != nullptr && temp_bytes >= s.bytes, "temporary buffer too small");
  IMP_REQUIRE(((uintptr_t)temp & 15) == 0, "temporary buffer must be 16-byte aligned");
  char* T = static_cast<char*>(temp);
  auto at = [&](size_t off) { return reinterpret_cast<u32*>(T + off); };
  const u32* pw = nullptr;
  if (OP == OP_HORNER) {
    hipLaunchKernelGGL(k_pow2_chain, dim3(1), dim3(1), 0, stream, base, at(s.pw), s.levels);
    IMP_CHECK(hipGetLastError());
  }
  for (u32 l = 0; l < s.levels; l++) {
    const View vin = l == 0 ? in : flat_view(at(s.sum[l - 1]), s.n[l]);
    if (OP == OP_HORNER) pw = at(s.pw) + 8 * SPW * l;
    u32* tsum = (l + 1 == s.levels && top_sum) ? top_sum : at(s.sum[l]);
    hipLaunchKernelGGL(k_scan_up<OP>, dim3((u32)s.tiles[l]), dim3(SB), 0, stream, vin, s.n[l],
                       at(s.lane[l]), tsum, pw);
    IMP_CHECK(hipGetLastError());
  }
  if (!down) return hipSuccess;
  for (int l = (int)s.levels - 1; l >= 0; l--) {
    const View vin = l == 0 ? in : flat_view(at(s.sum[l - 1]), s.n[l]);
    if (OP == OP_HORNER) pw = at(s.pw) + 8 * SPW * l;
    const bool top = l + 1 == (int)s.levels;
    const u32* c1 = top ? nullptr : at(s.carry[l]);
    const u32* c2 = (top || OP != OP_INV) ? nullptr : at(s.carry2[l]);
    if (l == 0) {
      hipLaunchKernelGGL((k_scan_down<OP, true>), dim3((u32)s.tiles[0]), dim3(SB), 0, stream,
                         vin, s.n[0], at(s.lane[0]), c1, c2, out, out, pw);
    } else {
      const View o1 = flat_view(at(s.carry[l - 1]), s.n[l]);
      const View o2 = OP == OP_INV ? flat_view(at(s.carry2[l - 1]), s.n[l]) : o1;
      hipLaunchKernelGGL((k_scan_down<OP, false>), dim3((u32)s.tiles[l]), dim3(SB), 0, stream,
                         vin, s.n[l], at(s.lane[l]), c1, c2, o1, o2, pw);
    }
    IMP_CHECK(hipGetLastError());
  }
  return hipSuccess;
}

// size query / size check of the reference's temp-buffer convention: with
// temp == NULL the size goes to *size and nothing runs
static bool size_query(void* temp, unsigned long* size, size_t need, hipError_t* err) {
  if (temp == nullptr) {
    if (size == nullptr) {
      *err = hipErrorInvalidValue;
      return true;
    }
    *size = (unsigned long)need;
    *err = hipSuccess;
    return true;
  }
  if (size != nullptr && *size < need) {
    fprintf(stderr, "[impeller_hip] invalid argument: temporary buffer of %lu bytes, need %zu\n",
            *size, need);
    *err = hipErrorInvalidValue;
    return true;
  }
  return false;
}

}  // namespace imp

using namespace imp;

static hipError_t binary_entry(int op, PolyPtr r, ConstPolyPtr a, ConstPolyPtr b,
                               hipStream_t stream) {
  if (a.len < b.len) {
    ConstPolyPtr t = a;
    a = b;
    b = t;
    if (op == EW_SUB) op = EW_RSUB;
  }
  const u64 len = b.len, upper = a.len;
  if (upper == 0) return hipSuccess;
  View va, vb, vr;
  IMP_REQUIRE(make_cview(a, &va) && make_view(r, &vr), "bad PolyPtr view");
  IMP_REQUIRE(len == 0 || make_cview(b, &vb), "bad PolyPtr view");
  if (len == 0) vb = va;
  IMP_REQUIRE(r.len >= upper, "result shorter than the longer operand");
  u32 blocks;
  IMP_CHECK(ew_grid(upper, &blocks));
  hipLaunchKernelGGL(k_poly_binary, dim3(blocks), dim3(256), 0, stream, op, va, vb, vr, len,
                     upper);
  return hipGetLastError();
}

// core/src/poly/src/poly.h:6-10 (poly_basic.cu:4-54)
extern "C" hipError_t poly_add(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream) {
  return binary_entry(EW_ADD, r, a, b, stream);
}
extern "C" hipError_t poly_sub(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream) {
  return binary_entry(EW_SUB, r, a, b, stream);
}
extern "C" hipError_t poly_mul(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream) {
  return binary_entry(EW_MUL, r, a, b, stream);
}

static hipError_t fill_entry(int mode, PolyPtr t, u64 s_begin, u64 s_len, hipStream_t stream) {
  if (t.len == 0) return hipSuccess;
  View v;
  IMP_REQUIRE(make_view(t, &v), "bad PolyPtr view");
  u32 blocks;
  IMP_CHECK(ew_grid(t.len, &blocks));
  hipLaunchKernelGGL(k_poly_fill, dim3(blocks), dim3(256), 0, stream, mode, v, (u64)t.len,
                     s_begin, s_len);
  return hipGetLastError();
}

// core/src/poly/src/poly.h:12-16 (poly_basic.cuh:84-120)
extern "C" hipError_t poly_one_lagrange(PolyPtr target, hipStream_t stream) {
  return fill_entry(1, target, 0, 0, stream);
}
extern "C" hipError_t poly_one_coef(PolyPtr target, hipStream_t stream,
                                   const PolySlice* slice_info) {
  if (slice_info == nullptr) return fill_entry(2, target, 0, target.len, stream);
  return fill_entry(2, target, slice_info->begin, slice_info->len, stream);
}
extern "C" hipError_t poly_zero(PolyPtr target, hipStream_t stream) {
  return fill_entry(0, target, 0, 0, stream);
}

// core/src/poly/src/poly.h:26-28 (poly_basic.cu:72-80)
extern "C" hipError_t inv_scalar(unsigned int* target, hipStream_t stream) {
  IMP_REQUIRE(target != nullptr, "null scalar");
  hipLaunchKernelGGL(k_inv_scalar, dim3(1), dim3(1), 0, stream, target);
  return hipGetLastError();
}
extern "C" hipError_t scalar_pow(unsigned int* target, unsigned long long exp,
                                 hipStream_t stream) {
  IMP_REQUIRE(target != nullptr, "null scalar");
  hipLaunchKernelGGL(k_scalar_pow, dim3(1), dim3(1), 0, stream, target, (u64)exp);
  return hipGetLastError();
}

// core/src/poly/src/poly.h:22 (scan_mul.cuh:6-26): in-place inclusive
// prefix products of target
extern "C" hipError_t scan_mul(void* temp_buffer, unsigned long* buffer_size, PolyPtr target,
                               hipStream_t stream) {
  const size_t need = scan_bytes(target.len, OP_PROD);
  hipError_t e;
  if (size_query(temp_buffer, buffer_size, need, &e)) return e;
  if (target.len == 0) return hipSuccess;
  View v;
  IMP_REQUIRE(make_view(target, &v), "bad PolyPtr view");
  return run_scan<OP_PROD>(v, target.len, v, temp_buffer, need, nullptr, true, nullptr, stream);
}

// core/src/poly/src/poly.h:24 (batched_invert.cuh:50-101): every non-zero
// element replaced by its inverse, zeros kept, in place
extern "C" hipError_t batched_invert(void* temp_buffer, unsigned long* buffer_size, PolyPtr poly,
                                     hipStream_t stream) {
  const size_t need = scan_bytes(poly.len, OP_INV);
  hipError_t e;
  if (size_query(temp_buffer, buffer_size, need, &e)) return e;
  if (poly.len == 0) return hipSuccess;
  View v;
  IMP_REQUIRE(make_view(poly, &v), "bad PolyPtr view");
  return run_scan<OP_INV>(v, poly.len, v, temp_buffer, need, nullptr, true, nullptr, stream);
}

// core/src/poly/src/poly.h:18 (poly_eval.cuh:7-71): *res = sum_i poly[i] x^i
// (x and res: one device element each)
extern "C" hipError_t poly_eval(void* temp_buf, unsigned long* temp_buf_size, ConstPolyPtr poly,
                                unsigned int* res, const unsigned int* x, hipStream_t stream) {
  const size_t need = scan_bytes(poly.len, OP_HORNER);
  hipError_t e;
  if (size_query(temp_buf, temp_buf_size, need, &e)) return e;
  IMP_REQUIRE(res != nullptr && x != nullptr, "null scalar");
  if (poly.len == 0) {
    hipLaunchKernelGGL(k_store_zero, dim3(1), dim3(1), 0, stream, res);
    return hipGetLastError();
  }
  View v;
  IMP_REQUIRE(make_cview(poly, &v), "bad PolyPtr view");
  return run_scan<OP_HORNER>(v, poly.len, v, temp_buf, need, x, false, res, stream);
}

// core/src/poly/src/poly.h:20 (kate_division.cuh:48-88): q = (p(X) - p(b)) /
// (X - b) for p of 2^log_p coefficients: q[i] = sum_{j > i} p[j] b^(j-i-1),
// q[2^log_p - 1] = 0 (halo2 kate_division padded with one zero)
extern "C" hipError_t kate_division(void* temp_buf, unsigned long* temp_buf_size,
                                    unsigned int log_p, ConstPolyPtr p, const unsigned int* b,
                                    PolyPtr q, hipStream_t stream) {
  IMP_REQUIRE(log_p < 63, "log_p out of range");
  const u64 len = 1ull << log_p;
  const size_t need = scan_bytes(len, OP_HORNER);
  hipError_t e;
  if (size_query(temp_buf, temp_buf_size, need, &e)) return e;
  IMP_REQUIRE(p.len == len && q.len == len, "p and q must hold 2^log_p elements");
  IMP_REQUIRE(b != nullptr, "null scalar");
  View vp, vq;
  IMP_REQUIRE(make_cview(p, &vp) && make_view(q, &vq), "bad PolyPtr view");
  return run_scan<OP_HORNER>(vp, len, vq, temp_buf, need, b, true, nullptr, stream);
}

// core/src/poly/src/poly.h:30 (permute.cuh): the lookup permutation of the
// first `usable` rows of input and table.  res_input / res_table hold usable
// elements.  The temp-buffer size depends on usable only.
extern "C" hipError_t permute(void* temp_buffer, unsigned long* buffer_size, unsigned long usable,
                              ConstPolyPtr input, ConstPolyPtr table, PolyPtr res_input,
                              PolyPtr res_table, hipStream_t stream) {
  PermPlan pp;
  perm_plan(usable, &pp);
  hipError_t e;
  if (size_query(temp_buffer, buffer_size, pp.bytes, &e)) return e;
  IMP_REQUIRE(input.len == table.len && usable <= input.len, "input and table lengths");
  IMP_REQUIRE(res_input.len == usable && res_table.len == usable, "result lengths");
  IMP_REQUIRE(((uintptr_t)temp_buffer & 15) == 0, "temporary buffer must be 16-byte aligned");
  if (usable == 0) return hipSuccess;
  View vi, vt, ri, rt;
  IMP_REQUIRE(make_cview(input, &vi) && make_cview(table, &vt) && make_view(res_input, &ri) &&
                  make_view(res_table, &rt),
              "bad PolyPtr view");
  char* T = static_cast<char*>(temp_buffer);
  auto at = [&](size_t off) { return reinterpret_cast<u32*>(T + off); };
  const u64 n = usable;
  const u32 g = (u32)((n + 255) / 256);
  // each column as numbers into k2, sorted from there into k0 (inputs) and
  // k1 (table)
  hipLaunchKernelGGL(k_to_number, dim3(g), dim3(256), 0, stream, vi, n, at(pp.k2));
  IMP_CHECK(hipGetLastError());
  IMP_CHECK(sort256(at(pp.k0), at(pp.k2), n, T + pp.sort, stream));
  hipLaunchKernelGGL(k_to_number, dim3(g), dim3(256), 0, stream, vt, n, at(pp.k2));
  IMP_CHECK(hipGetLastError());
  IMP_CHECK(sort256(at(pp.k1), at(pp.k2), n, T + pp.sort, stream));
  hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, stream, at(pp.avail), n, 1u);
  IMP_CHECK(hipGetLastError());
  // the table windows of the input blocks, in the pos area (2 words per 256
  // inputs; pos is first written after the lookup)
  hipLaunchKernelGGL(k_perm_window, dim3((u32)((2 * ((n + 255) / 256) + 255) / 256)), dim3(256),
                     0, stream, at(pp.k0), at(pp.k1), n, at(pp.pos));
  IMP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_perm_lookup, dim3(g), dim3(256), 0, stream, at(pp.k0), at(pp.k1), n,
                     at(pp.pos), at(pp.start), at(pp.avail), rt);
  IMP_CHECK(hipGetLastError());
  // positions of the leftover table values, then the run starts before each row
  IMP_CHECK(hipMemcpyAsync(at(pp.pos), at(pp.avail), n * 4, hipMemcpyDeviceToDevice, stream));
  IMP_CHECK(u32scan(at(pp.pos), n, T + pp.scan, stream));
  hipLaunchKernelGGL(k_perm_compact, dim3(g), dim3(256), 0, stream, at(pp.k1), at(pp.avail),
                     at(pp.pos), n, at(pp.k2));
  IMP_CHECK(hipGetLastError());
  IMP_CHECK(u32scan(at(pp.start), n, T + pp.scan, stream));
  hipLaunchKernelGGL(k_perm_fill, dim3(g), dim3(256), 0, stream, at(pp.k0), at(pp.k2),
                     at(pp.start), n, ri, rt);
  return hipGetLastError();
}
