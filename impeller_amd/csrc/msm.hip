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
//                     (8-bit digits, per-tile histograms, two stable 4-bit
//                     splits per tile in LDS, coalesced run writes)
//   k_bucket_bounds   run boundaries of each bucket in the sorted keys
//   kb_scan_a/b/c     device-wide segmented exclusive scans (chunk sums,
//                     chunk offsets, apply) for the radix histograms, the
//                     task offsets and the combine levels
//   k_bucket_tasks    each bucket split into tasks of <= T entries, one
//                     thread per task, mixed XYZZ + affine additions
//   k_bucket_combine  per-bucket partial-sum tree (groups of 2^g); each level
//                     runs one thread per live group (compact indices from a
//                     scan), so a heavily loaded bucket never serialises on one
//                     thread and idle buckets cost no lanes
//   k_msm_reduce      running sums over segments of 4 buckets
//   k_msm_sum         segment sums split by segment-index bit
//   k_msm_final       doubling tree per window -> Horner over windows -> affine
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "field29.h"
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
  u32 passes;     // radix passes (8-bit digits over c-bit keys)
  u32 T;          // entries per accumulation task
  u32 g;          // log2 of the combine group size
  u32 levels;     // combine levels (worst case over bucket loads)
  u32 chunks;     // scan chunks per window (max over the scans)
  u64 task_cap;   // task slots per physical window
  u32 seg;        // buckets per reduce segment (4, or 8 for c >= 17)
  u32 nseg;       // reduce segments per physical window (nb / seg)
  u32 lbits;      // log2(nseg)
  u32 sbits;      // log2(seg)
  u32 g0;         // level-0 work-groups of the sum T_s (power of two, 1024 inputs each)
  u32 sum_levels; // 1: the final reduces the g0 partials; 2: a second level does
  u32 fin_g;      // partials per window sum that the final reads (g0 or 1)
};

static const u32 kSortTile = 4096;  // 256 threads x 16 keys
static const u32 kPt = 36;           // u32 per stored XYZZ point (4 x 9 limbs of 29 bits)
static const u32 kRadixBits = 8;     // digit bits per sort pass
static const u32 kRadix = 1u << kRadixBits;

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
  s.passes = (c + kRadixBits - 1) / kRadixBits;  // keys <= 2^(c-1) need c bits
  s.T = task;
  s.g = 4;
  s.task_cap = s.seg_len / s.T + s.nb + 1;
  u64 max_tasks = s.seg_len / s.T + 1;
  s.levels = 0;
  while ((1ull << (s.g * s.levels)) < max_tasks && s.levels < 8) s.levels++;
  u64 hl = (u64)kRadix * s.tiles, bl = (u64)s.nb + 1;
  s.chunks = (u32)(((hl > bl ? hl : bl) + 4095) / 4096);
  // longer segments for wide windows: halves the bit-split sums' work for a
  // few more dependent additions per segment
  s.sbits = c >= 17 ? 3 : 2;
  s.seg = 1u << s.sbits;
  s.nseg = s.nb / s.seg;  // nb >= 8 for c >= 4
  s.lbits = c - 1 - s.sbits;
  // window sums: lbits + 1 of them (the T_s over all nseg segments, and one
  // per segment-index bit over the nseg / 2 segments with that bit set);
  // level 0 gives each work-group 1024 inputs (four per thread, then a
  // tree), and when the final block cannot hold all g0 partial slots of
  // every sum a second level reduces them to one
  const u32 nv = s.lbits + s.sbits;  // = c - 1
  s.g0 = s.nseg > 1024 ? s.nseg / 1024 : 1;
  s.sum_levels = (u64)s.n_windows * nv * s.g0 <= 256 ? 1 : 2;
  s.fin_g = s.sum_levels == 1 ? s.g0 : 1;
  return s;
}

struct MsmLayout {
  size_t keysA, valsA, keysB, valsB, hist, bstart, bend, csum, toff, loff, partA, partB, rparts,
      sparts, sparts1, maxnt, conv, total;
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
  l.hist = o; o = align256(o + (size_t)s.n_windows * kRadix * s.tiles * 4 + 4 * s.n_windows);
  l.bstart = o; o = align256(o + (size_t)s.n_windows * (s.nb + 1) * 4);
  l.bend = o; o = align256(o + (size_t)s.n_windows * (s.nb + 1) * 4);
  l.csum = o; o = align256(o + (size_t)s.n_windows * (s.chunks + 1) * 4);
  l.toff = o; o = align256(o + (size_t)s.n_windows * (s.nb + 2) * 4);
  l.loff = o; o = align256(o + (size_t)s.n_windows * (s.nb + 2) * 4);
  l.partA = o; o = align256(o + (size_t)s.n_windows * s.task_cap * kPt * 4);
  l.partB = o; o = align256(o + (size_t)s.n_windows * s.task_cap * kPt * 4);
  l.rparts = o; o = align256(o + (size_t)s.n_windows * 2 * s.nseg * kPt * 4);
  l.sparts = o; o = align256(o + (size_t)s.n_windows * (s.lbits + 1) * s.g0 * kPt * 4);
  l.sparts1 = o; o = align256(o + (size_t)s.n_windows * (s.lbits + 1) * kPt * 4);
  l.maxnt = o; o = align256(o + (size_t)s.n_windows * 4);
  // one table (the plain bases, affine R-form) is converted per call
  l.conv = o; o = align256(o + (s.n_pre == 1 ? (size_t)s.N * 64 : 0));
  l.total = o;
  return l;
}

// ----------------------------------------------------------------- job
// One MSM problem: its shape plus every device pointer it needs.  The same
// kernels run a single MSM (a batch of one) or many uneven MSMs at once.
struct MsmJob {
  MsmShape s;
  u64 tstride;  // points between consecutive tables (>= N)
  const u32* bases;  // point tables in the table format
  const u32* raw;    // one table of affine R-form bases to convert into bases, or null
  const u32* scalars;
  u32* result;
  u32 *kA, *vA, *kB, *vB, *hist, *bstart, *bend, *csum, *toff, *loff, *partA, *partB, *rparts,
      *sparts, *sparts1;
  // the most tasks any bucket of a window has (set by the task-offset scan):
  // combine level x has work only where some bucket has more than
  // 2^(g (x - 1)) tasks, and its scans and combine return at once otherwise
  u32* maxnt;
  // bucket sums kept across the parts of one MSM (n_windows x nb XYZZ, or
  // null): the accumulate stage stores (bmode 1, the first part) or adds
  // (bmode 2) this part's bucket totals, and the reduce stage of the last
  // part reads them instead of this part's partial-sum tree
  u32* bacc;
  u32 bmode;
};

static MsmJob make_msm_job(const MsmShape& s, u64 tstride, const u32* bases, const u32* scalars,
                           u32* result, char* ws) {
  MsmLayout l = msm_layout(s);
  MsmJob J;
  J.s = s;
  J.tstride = tstride;
  J.bases = bases;
  J.raw = nullptr;
  if (s.n_pre == 1) {
    J.raw = bases;
    J.bases = (const u32*)(ws + l.conv);
  }
  J.scalars = scalars;
  J.result = result;
  J.kA = (u32*)(ws + l.keysA);
  J.vA = (u32*)(ws + l.valsA);
  J.kB = (u32*)(ws + l.keysB);
  J.vB = (u32*)(ws + l.valsB);
  J.hist = (u32*)(ws + l.hist);
  J.bstart = (u32*)(ws + l.bstart);
  J.bend = (u32*)(ws + l.bend);
  J.csum = (u32*)(ws + l.csum);
  J.toff = (u32*)(ws + l.toff);
  J.loff = (u32*)(ws + l.loff);
  J.partA = (u32*)(ws + l.partA);
  J.partB = (u32*)(ws + l.partB);
  J.rparts = (u32*)(ws + l.rparts);
  J.sparts = (u32*)(ws + l.sparts);
  J.sparts1 = (u32*)(ws + l.sparts1);
  J.maxnt = (u32*)(ws + l.maxnt);
  J.bacc = nullptr;
  J.bmode = 0;
  return J;
}

__device__ __forceinline__ const u32* sorted_keys(const MsmJob& J) {
  return (J.s.passes & 1) ? J.kB : J.kA;
}
__device__ __forceinline__ const u32* sorted_vals(const MsmJob& J) {
  return (J.s.passes & 1) ? J.vB : J.vA;
}
// combine levels bucket b needs: smallest l with 2^(g l) >= nt (0 for nt <= 1);
// its total ends in partA for even l, partB for odd l
// The accumulation chunks of a window: T consecutive sorted entries each,
// counted from z, the number of zero digits (bucket 0, sorted first and never
// added).  The tasks of a bucket b >= 1 whose entries are [bs, be) are the
// chunks that meet it (0 for an empty bucket); over a window they number at
// most ceil(seg_len / T) + nb - 1 < task_cap.
__host__ __device__ __forceinline__ u32 bucket_tasks(u32 bs, u32 be, u32 z, u32 T) {
  return be > bs ? (be - 1 - z) / T - (bs - z) / T + 1 : 0u;
}
__host__ __device__ inline u64 chunks_per_window(const MsmShape& s) {
  return (s.seg_len + s.T - 1) / s.T;
}

__device__ __forceinline__ u32 bucket_levels(u32 nt, u32 g) {
  u32 l = 0;
  while ((1ull << (g * l)) < nt) l++;
  return l;
}

// ------------------------------------------------------------ stage bodies
// Signed c-bit digits (msm_impl.cuh:22-33; a digit equal to 2^(c-1) stays
// positive, so the top window never carries out).  Logical window w goes to
// physical window w % n_windows and table w / n_windows (msm_impl.cuh:56-67).
// key = bucket |digit|, val = point index (table * tstride + i) | sign << 31.
__device__ __forceinline__ bool aligned16(const u32* p) { return ((uintptr_t)p & 15) == 0; }

// Signed c-bit digits of the scalars of points i0 .. i0 + 3 (thread-wise):
// each scalar (as an integer) is shifted right by c after every window, so
// window w's bits are always the low bits of limb 0 (no indexing into the
// limbs by a run-time position); each slot's four keys and four values go
// out as one 16-byte store each when aligned.
__device__ __forceinline__ void digits_body(const MsmJob& J, u64 i0) {
  const MsmShape& s = J.s;
  if (i0 >= s.N) return;
  const u32 np = s.N - i0 < 4 ? (u32)(s.N - i0) : 4u;
  u32 v[4][8];
#pragma unroll
  for (u32 p = 0; p < 4; p++) {
    const Fr x = p < np ? to_number<FrParams>(load_fe<Fr>(J.scalars + 8 * (i0 + p))) : Fr::zero();
#pragma unroll
    for (u32 l = 0; l < 8; l++) v[p][l] = x.v[l];
  }
  const u32 c = s.c, half = 1u << (c - 1), mask = (1u << c) - 1;
  u32 carry[4] = {0, 0, 0, 0};
  const u32 slots = s.n_windows * s.n_pre;
  for (u32 w = 0; w < slots; w++) {
    const u32 pw = w % s.n_windows, grp = w / s.n_windows;
    const u64 idx = (u64)pw * s.seg_len + (u64)grp * s.N + i0;
    u32 key[4], val[4];
#pragma unroll
    for (u32 p = 0; p < 4; p++) {
      u32 k = 0, sign = 0;
      if (w < s.W) {
        const u32 d = (v[p][0] & mask) + carry[p];
        if (d > half) {
          k = (1u << c) - d;
          sign = 0x80000000u;
          carry[p] = 1;
        } else {
          k = d;
          carry[p] = 0;
        }
#pragma unroll
        for (u32 l = 0; l < 7; l++) v[p][l] = (v[p][l] >> c) | (v[p][l + 1] << (32 - c));
        v[p][7] >>= c;
      }
      key[p] = k;
      val[p] = (u32)((u64)grp * J.tstride + i0 + p) | sign;
    }
    if (np == 4 && aligned16(J.kA + idx) && aligned16(J.vA + idx)) {
      gst(reinterpret_cast<uint4*>(J.kA + idx), make_uint4(key[0], key[1], key[2], key[3]));
      gst(reinterpret_cast<uint4*>(J.vA + idx), make_uint4(val[0], val[1], val[2], val[3]));
    } else {
#pragma unroll
      for (u32 p = 0; p < 4; p++) {
        if (p < np) {
          gst(J.kA + idx + p, key[p]);
          gst(J.vA + idx + p, val[p]);
        }
      }
    }
  }
}

// one affine R-form base -> the table format
__device__ __forceinline__ void convert_body(const MsmJob& J, u64 i) {
  if (i >= J.s.N || J.raw == nullptr) return;
  aff_to_table(const_cast<u32*>(J.bases) + 16 * i, load_aff(J.raw + 16 * i));
}

// 16 consecutive words from p (16-byte aligned) in four 16-byte loads: a
// thread's run of 16 keys or counts, read with a quarter of the instructions
// that word loads would take (lanes 64 bytes apart).
__device__ __forceinline__ void gld16(const u32* p, u32 v[16]) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint4 w = gld(q + k);
    v[4 * k] = w.x;
    v[4 * k + 1] = w.y;
    v[4 * k + 2] = w.z;
    v[4 * k + 3] = w.w;
  }
}

// LSD radix sort, 8-bit digits, stable per pass.  Tile histograms go to
// hist[(pw * 256 + digit) * tiles + tile]; their exclusive scan in that order
// is each (digit, tile) run's global start.
// Digit counts of one tile: thread t counts its run of 16 keys [16 t,
// 16 t + 16), read with four 16-byte loads, into its wave's own 256 counters
// (atomics of different waves never meet), and the four waves' counters are
// summed per digit at the end.  cnt holds 4 x 256 words.
__device__ __forceinline__ void hist_body(const MsmJob& J, u32 pass, u64 lb, u32* cnt) {
  const MsmShape& s = J.s;
  const u32 tile = (u32)(lb % s.tiles), pw = (u32)(lb / s.tiles), t = threadIdx.x;
  const u32 shift = kRadixBits * pass;
  const u32* keys = (pass & 1) ? J.kB : J.kA;
  __syncthreads();
  for (u32 b = t; b < 4 * kRadix; b += 256) cnt[b] = 0;
  __syncthreads();
  const u32* kk = keys + (u64)pw * s.seg_len;
  const u64 i0 = (u64)tile * kSortTile + 16 * t;
  const u32 n = i0 < s.seg_len ? (s.seg_len - i0 < 16 ? (u32)(s.seg_len - i0) : 16u) : 0u;
  u32 k[16];
  if (n == 16 && aligned16(kk)) {
    gld16(kk + i0, k);
  } else {
#pragma unroll
    for (u32 j = 0; j < 16; j++) k[j] = j < n ? gld(kk + i0 + j) : 0u;
  }
  u32* mine = cnt + (t >> 6) * kRadix;
#pragma unroll
  for (u32 j = 0; j < 16; j++)
    if (j < n) atomicAdd(&mine[(k[j] >> shift) & (kRadix - 1)], 1u);
  __syncthreads();
  for (u32 b = t; b < kRadix; b += 256)
    gst(J.hist + ((u64)pw * kRadix + b) * s.tiles + tile,
        cnt[b] + cnt[kRadix + b] + cnt[2 * kRadix + b] + cnt[3 * kRadix + b]);
}

__device__ __forceinline__ u32 block_excl_scan(u32 v, u32* part, u32* total);

// block-wide exclusive scan of len entries (in place allowed; blockDim.x ==
// 256); with write_total, dst[len] = total
__device__ __forceinline__ void scan_body(const u32* src, u32* dst, u64 len, bool write_total, u32* part) {
  const u32 T = 256, t = threadIdx.x;
  const u64 per = (len + T - 1) / T;
  u64 lo = (u64)t * per, hi = lo + per < len ? lo + per : len;
  u32 sum = 0;
  for (u64 b = lo; b < hi; b++) sum += src[b];
  u32 total;
  u32 run = block_excl_scan(sum, part, &total);
  for (u64 b = lo; b < hi; b++) {
    u32 x = src[b];
    dst[b] = run;
    run += x;
  }
  if (write_total && t == T - 1) dst[len] = total;
}

// exclusive scan of one value per thread (blockDim.x == 256, four waves of
// 64); returns the thread's prefix, *total the block sum.  Each wave scans
// its lanes with cross-lane reads (six steps, no barrier); the four wave
// totals meet in part[0..3] behind one barrier.
__device__ __forceinline__ u32 block_excl_scan(u32 v, u32* part, u32* total) {
  const u32 t = threadIdx.x, lane = t & 63, wv = t >> 6;
  u32 inc = v;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    const u32 w = __shfl(inc, (int)(lane >= d ? lane - d : lane), 64);
    if (lane >= d) inc += w;
  }
  __syncthreads();  // the caller's previous reads of part are done
  if (lane == 63) part[wv] = inc;
  __syncthreads();
  const u32 p0 = part[0], p1 = part[1], p2 = part[2], p3 = part[3];
  *total = p0 + p1 + p2 + p3;
  const u32 off = (wv > 0 ? p0 : 0u) + (wv > 1 ? p1 : 0u) + (wv > 2 ? p2 : 0u);
  return off + inc - v;
}

struct ScatterLds {
  union {
    u32 cnt[16 * 256];  // splits: per (field value, thread) counts, then running ranks (swizzled)
    struct {
      u32 gbase[kRadix];  // scatter: global start of each digit's run of this tile
      u32 dbase[kRadix];  // first staged position of each digit
    } run;
  };
  u32 part[256];
  u32 sk[kSortTile];  // the tile in (digit, input order), swizzled: staged for
  u32 sv[kSortTile];  // coalesced writes
};

// LDS word swizzle for the 4096-word arrays: thread t's run of 16 words
// [16t, 16t + 16) lands in 64 distinct banks across a wave (the low four
// address bits are XORed with bits 6..9), while 64 consecutive words stay in
// 64 distinct banks
__device__ __forceinline__ u32 swz(u32 i) { return i ^ ((i >> 6) & 15u); }

// Stable split of the tile by the field of `bits` (<= 4) bits at bit sh.
// Thread t holds tile positions 16t .. 16t + n - 1 in key/val; ranks come
// from per-thread field counts scanned in (field, thread) order, and the
// entries land in sk/sv.
__device__ __forceinline__ void split_field(ScatterLds& L, const u32 key[16], const u32 val[16],
                                            u32 n, u32 sh, u32 bits) {
  const u32 mask = (1u << bits) - 1;
  const u32 t = threadIdx.x;
#pragma unroll
  for (u32 d = 0; d < 16; d++) L.cnt[swz(d * 256 + t)] = 0;
#pragma unroll
  for (u32 j = 0; j < 16; j++)
    if (j < n) L.cnt[swz(((key[j] >> sh) & mask) * 256 + t)] += 1;
  __syncthreads();
  // thread t scans the flattened counts [16t, 16t + 16)
  u32 loc[16];
  u32 sum = 0;
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    loc[j] = L.cnt[swz(t * 16 + j)];
    sum += loc[j];
  }
  u32 total;
  u32 run = block_excl_scan(sum, L.part, &total);
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    L.cnt[swz(t * 16 + j)] = run;
    run += loc[j];
  }
  __syncthreads();
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    if (j < n) {
      const u32 c = swz(((key[j] >> sh) & mask) * 256 + t);
      const u32 r = L.cnt[c];
      L.cnt[c] = r + 1;
      L.sk[swz(r)] = key[j];
      L.sv[swz(r)] = val[j];
    }
  }
  __syncthreads();
}

// One tile of 4096 entries: stable 4-bit splits order it by the digit
// (input order kept within a digit; one split for a top digit of at most 4
// bits), staged in LDS; then consecutive threads write consecutive positions
// of each digit's run: staged entry i of digit d goes to
// gbase[d] + (i - dbase[d]).  (10-bit digits, two passes instead of three at
// c = 20, measured no faster: runs of 4 entries per digit and tile.)
__device__ __forceinline__ void scatter_body(const MsmJob& J, u32 pass, u64 lb, ScatterLds& L) {
  const MsmShape& s = J.s;
  const u32 tile = (u32)(lb % s.tiles), pw = (u32)(lb / s.tiles), t = threadIdx.x;
  const u32 shift = kRadixBits * pass;
  const u32* ksrc = (pass & 1) ? J.kB : J.kA;
  const u32* vsrc = (pass & 1) ? J.vB : J.vA;
  u32* kdst = (pass & 1) ? J.kA : J.kB;
  u32* vdst = (pass & 1) ? J.vA : J.vB;
  const u64 seg0 = (u64)pw * s.seg_len;
  const u64 tile0 = (u64)tile * kSortTile;
  const u32 n_tile = (u32)(s.seg_len - tile0 < kSortTile ? s.seg_len - tile0 : kSortTile);
  const u32 n = n_tile > 16 * t ? (n_tile - 16 * t < 16 ? n_tile - 16 * t : 16) : 0;
  u32 key[16], val[16];
  __syncthreads();
  if (n == 16 && aligned16(ksrc + seg0) && aligned16(vsrc + seg0)) {
    gld16(ksrc + seg0 + tile0 + 16 * t, key);
    gld16(vsrc + seg0 + tile0 + 16 * t, val);
  } else {
#pragma unroll
    for (u32 j = 0; j < 16; j++) {
      if (j < n) {
        key[j] = gld(ksrc + seg0 + tile0 + 16 * t + j);
        val[j] = gld(vsrc + seg0 + tile0 + 16 * t + j);
      }
    }
  }
  // keys are below 2^c: the top digit may be narrower
  const u32 wbits = s.c - shift < kRadixBits ? s.c - shift : kRadixBits;
  for (u32 b = 0; b < wbits; b += 4) {
    if (b) {
#pragma unroll
      for (u32 j = 0; j < 16; j++) {
        if (j < n) {
          key[j] = L.sk[swz(16 * t + j)];
          val[j] = L.sv[swz(16 * t + j)];
        }
      }
      __syncthreads();
    }
    split_field(L, key, val, n, shift + b, wbits - b < 4 ? wbits - b : 4);
  }
  // digit run starts in the staged order, and the runs' global starts (the
  // split counters are dead: gbase / dbase reuse their LDS)
  for (u32 b = t; b < kRadix; b += 256)
    L.run.gbase[b] = gld(J.hist + ((u64)pw * kRadix + b) * s.tiles + tile);
  for (u32 j = 0; j < n; j++) {
    const u32 i = 16 * t + j;
    const u32 d = (L.sk[swz(i)] >> shift) & (kRadix - 1);
    if (i == 0 || ((L.sk[swz(i - 1)] >> shift) & (kRadix - 1)) != d) L.run.dbase[d] = i;
  }
  __syncthreads();
  for (u32 i = t; i < n_tile; i += blockDim.x) {
    const u32 k = L.sk[swz(i)];
    const u32 d = (k >> shift) & (kRadix - 1);
    const u64 pos = (u64)L.run.gbase[d] + i - L.run.dbase[d];
    gst(kdst + seg0 + pos, k);
    gst(vdst + seg0 + pos, L.sv[swz(i)]);
  }
}

__device__ __forceinline__ void zero_body(const MsmJob& J, u64 i) {
  u64 nbw = (u64)J.s.n_windows * (J.s.nb + 1);
  if (i < nbw) J.bstart[i] = 0;
  else if (i < 2 * nbw) J.bend[i - nbw] = 0;
  else if (i < 2 * nbw + J.s.n_windows) J.maxnt[i - 2 * nbw] = 0;
}

// bucket run bounds in the sorted keys (bucket 0, the zero digits, first:
// its end is their count): work-group lb covers 4096 consecutive keys of one
// physical window, thread t the run of 16 keys [16 t, 16 t + 16) of it, read
// with four 16-byte loads plus the keys just before and after the run
static const u32 kBoundsSpan = 4096;

__host__ __device__ inline u64 bounds_blocks_per_window(const MsmShape& s) {
  return (s.seg_len + kBoundsSpan - 1) / kBoundsSpan;
}

__device__ __forceinline__ void bounds_body(const MsmJob& J, u64 lb) {
  const MsmShape& s = J.s;
  const u64 bpw = bounds_blocks_per_window(s);
  const u32 pw = (u32)(lb / bpw);
  const u64 i0 = (lb - (u64)pw * bpw) * kBoundsSpan + 16 * threadIdx.x;
  if (pw >= s.n_windows || i0 >= s.seg_len) return;
  const u32* kk = sorted_keys(J) + (u64)pw * s.seg_len;
  u32* bs = J.bstart + (u64)pw * (s.nb + 1);
  u32* be = J.bend + (u64)pw * (s.nb + 1);
  const u32 n = s.seg_len - i0 < 16 ? (u32)(s.seg_len - i0) : 16u;
  u32 k[16];
  if (n == 16 && aligned16(kk)) {
    gld16(kk + i0, k);
  } else {
#pragma unroll
    for (u32 j = 0; j < 16; j++) k[j] = j < n ? gld(kk + i0 + j) : 0u;
  }
  const u32 before = i0 ? gld(kk + i0 - 1) : 0xffffffffu;
  const u32 after = i0 + n < s.seg_len ? gld(kk + i0 + n) : 0xffffffffu;
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    if (j < n) {
      const u32 prev = j ? k[j - 1] : before;
      const u32 next = j + 1 < n ? k[j + 1] : after;
      if (k[j] != prev) gst(bs + k[j], (u32)(i0 + j));
      if (k[j] != next) gst(be + k[j], (u32)(i0 + j + 1));
    }
  }
}

// ----------------------------------------------------------- segmented scans
// One exclusive scan per physical window over one of three inputs:
//   SC_HIST   the radix histograms hist[(pw * 256 + digit) * tiles + tile], in place;
//   SC_TASKS  tasks per bucket, the chunks meeting it (0 for bucket 0) -> toff, total at [nb+1];
//   SC_LEVEL  groups of combine level x per bucket (0 for a bucket already reduced
//             to one entry) -> loff, total at [nb+1].
// Phase A sums chunks of 4096 inputs, phase B scans the chunk sums of a window
// (and stores the total), phase C scans each chunk and adds its offset.
enum ScanKind : u32 { SC_HIST = 0, SC_TASKS = 1, SC_LEVEL = 2 };
static const u32 kScanChunk = 4096;  // 256 threads x 16 inputs

__host__ __device__ inline u64 scan_len(const MsmShape& s, u32 kind) {
  return kind == SC_HIST ? (u64)kRadix * s.tiles : (u64)s.nb + 1;
}
__host__ __device__ inline u32 scan_chunks(const MsmShape& s, u32 kind) {
  return (u32)((scan_len(s, kind) + kScanChunk - 1) / kScanChunk);
}

// groups of combine level x of a bucket with nt tasks (0 once one is left)
__device__ __forceinline__ u32 level_groups(u32 nt, u32 x, u32 g) {
  if (x == 0) return nt;  // level 0: the tasks themselves
  const u32 shp = g * (x - 1), sh = g * x;
  const u32 nprev = (u32)(((u64)nt + (1ull << shp) - 1) >> shp);
  return nprev > 1 ? (u32)(((u64)nt + (1ull << sh) - 1) >> sh) : 0u;
}

// combine level x (>= 1) of window pw has no bucket to work on: no bucket
// has more than 2^(g (x - 1)) tasks (level_groups is then 0 everywhere)
__device__ __forceinline__ bool level_idle(const MsmJob& J, u32 x, u32 pw) {
  return gld(J.maxnt + pw) <= (1u << (J.s.g * (x - 1)));
}

__device__ __forceinline__ u32 scan_in(const MsmJob& J, u32 kind, u32 x, u32 pw, u64 i) {
  const MsmShape& s = J.s;
  if (kind == SC_HIST) return J.hist[(u64)pw * kRadix * s.tiles + i];
  if (i == 0) return 0;
  const u64 bo = (u64)pw * (s.nb + 1) + i;
  const u32 nt = bucket_tasks(J.bstart[bo], J.bend[bo], J.bend[(u64)pw * (s.nb + 1)], s.T);
  return kind == SC_TASKS ? nt : level_groups(nt, x, s.g);
}

// inputs base .. base + 15 of a scan: 16-byte loads when the run is whole and
// aligned, one value at a time otherwise
__device__ __forceinline__ void scan_vals16(const MsmJob& J, u32 kind, u32 x, u32 pw, u64 base,
                                            u64 len, u32 v[16]) {
  const MsmShape& s = J.s;
  if (base + 16 <= len) {
    if (kind == SC_HIST) {
      const u32* h = J.hist + (u64)pw * kRadix * s.tiles + base;
      if (aligned16(h)) {
        gld16(h, v);
        return;
      }
    } else {
      const u32* bs = J.bstart + (u64)pw * (s.nb + 1) + base;
      const u32* be = J.bend + (u64)pw * (s.nb + 1) + base;
      if (aligned16(bs) && aligned16(be)) {
        u32 a[16], b[16];
        gld16(bs, a);
        gld16(be, b);
        const u32 z = gld(J.bend + (u64)pw * (s.nb + 1));
#pragma unroll
        for (u32 j = 0; j < 16; j++) {
          const u32 nt = base + j == 0 ? 0u : bucket_tasks(a[j], b[j], z, s.T);
          v[j] = kind == SC_TASKS ? nt : level_groups(nt, x, s.g);
        }
        return;
      }
    }
  }
#pragma unroll
  for (u32 j = 0; j < 16; j++) v[j] = base + j < len ? scan_in(J, kind, x, pw, base + j) : 0u;
}

__device__ __forceinline__ u32* scan_out(const MsmJob& J, u32 kind, u32 pw) {
  if (kind == SC_HIST) return J.hist + (u64)pw * kRadix * J.s.tiles;
  return (kind == SC_TASKS ? J.toff : J.loff) + (u64)pw * (J.s.nb + 2);
}

__device__ __forceinline__ void scan_a_body(const MsmJob& J, u32 kind, u32 x, u64 lb, u32* part) {
  const u32 nch = scan_chunks(J.s, kind);
  const u32 pw = (u32)(lb / nch), ch = (u32)(lb % nch);
  if (kind == SC_LEVEL && level_idle(J, x, pw)) return;  // uniform over the work-group
  const u64 len = scan_len(J.s, kind);
  const u64 base = (u64)ch * kScanChunk + (u64)threadIdx.x * 16;
  u32 v[16];
  scan_vals16(J, kind, x, pw, base, len, v);
  u32 sum = 0, mx = 0;
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    sum += v[j];
    mx = v[j] > mx ? v[j] : mx;
  }
  if (kind == SC_TASKS) {
    // the window's largest task count: a wave maximum, one atomic per wave
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      const u32 o = (u32)__shfl_xor((int)mx, d);
      mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0 && mx > 1) atomicMax(J.maxnt + pw, mx);
  }
  u32 total;
  block_excl_scan(sum, part, &total);
  if (threadIdx.x == 0) J.csum[(u64)pw * (J.s.chunks + 1) + ch] = total;
}

__device__ __forceinline__ void scan_b_body(const MsmJob& J, u32 kind, u32 x, u64 pw, u32* part) {
  if (kind == SC_LEVEL && level_idle(J, x, (u32)pw)) return;
  const u32 nch = scan_chunks(J.s, kind);
  u32* cs = J.csum + pw * (J.s.chunks + 1);
  scan_body(cs, cs, nch, true, part);
  __syncthreads();
  if (threadIdx.x == 0 && kind != SC_HIST) scan_out(J, kind, (u32)pw)[scan_len(J.s, kind)] = cs[nch];
}

__device__ __forceinline__ void scan_c_body(const MsmJob& J, u32 kind, u32 x, u64 lb, u32* part) {
  const u32 nch = scan_chunks(J.s, kind);
  const u32 pw = (u32)(lb / nch), ch = (u32)(lb % nch);
  if (kind == SC_LEVEL && level_idle(J, x, pw)) return;
  const u64 len = scan_len(J.s, kind);
  const u64 base = (u64)ch * kScanChunk + (u64)threadIdx.x * 16;
  u32 v[16];
  scan_vals16(J, kind, x, pw, base, len, v);
  u32 sum = 0;
#pragma unroll
  for (u32 j = 0; j < 16; j++) sum += v[j];
  u32 total;
  u32 run = block_excl_scan(sum, part, &total) + J.csum[(u64)pw * (J.s.chunks + 1) + ch];
  u32* out = scan_out(J, kind, pw) + base;
  u32 o[16];
#pragma unroll
  for (u32 j = 0; j < 16; j++) {
    o[j] = run;
    run += v[j];
  }
  if (base + 16 <= len && aligned16(out)) {
    uint4* q = reinterpret_cast<uint4*>(out);
#pragma unroll
    for (u32 k = 0; k < 4; k++) gst(q + k, make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]));
  } else {
#pragma unroll
    for (u32 j = 0; j < 16; j++)
      if (base + j < len) out[j] = o[j];
  }
}

// largest b in [1, nb] with toff[b] <= t (toff nondecreasing over [1, nb],
// toff[nb + 1] its total).  First guess b = t + 1, exact when every bucket
// up to there has one group (the common case at combine level 1, where a
// bucket's few tasks form one group): two loads instead of a binary search
// of log2(nb) dependent loads.
__device__ __forceinline__ u32 find_bucket(const u32* toff, u32 nb, u32 t) {
  if (t + 1 <= nb && toff[t + 1] <= t && (t + 1 == nb || toff[t + 2] > t)) return t + 1;
  u32 lo = 1, hi = nb;
  while (lo < hi) {
    u32 mid = (lo + hi + 1) >> 1;
    if (toff[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// A sorted entry's table point as stored (x then y, 8 words each) and its
// sign: loaded one entry ahead, unpacked to limbs only when it is added.
typedef u32 u32x4 __attribute__((ext_vector_type(4)));
struct RawEntry {
  u32x4 q[4];
  u32 neg;
};

// Four aligned 16-byte global loads of one 64-byte point.  Left to itself
// the compiler narrows them to the words each 29-bit limb needs (seven
// overlapping flat loads at byte offsets 0/4/8/24/36/40/56: 46 % more L1
// accesses); unpack_entry therefore passes the four vectors through an empty
// asm statement, which uses every bit of them.
__device__ __forceinline__ RawEntry load_entry(const MsmJob& J, u32 e) {
  const u32x4* q = reinterpret_cast<const u32x4*>(J.bases + 16 * (u64)(e & 0x7fffffffu));
  RawEntry r;
#pragma unroll
  for (int k = 0; k < 4; k++) r.q[k] = gld(q + k);
  r.neg = e >> 31;
  return r;
}

__device__ __forceinline__ G1Aff29 unpack_entry(RawEntry r) {
#if !defined(IMP_EMU) && defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(r.q[0]), "+v"(r.q[1]), "+v"(r.q[2]), "+v"(r.q[3]));
#endif
  const u32 x[8] = {r.q[0].x, r.q[0].y, r.q[0].z, r.q[0].w, r.q[1].x, r.q[1].y, r.q[1].z, r.q[1].w};
  const u32 y[8] = {r.q[2].x, r.q[2].y, r.q[2].z, r.q[2].w, r.q[3].x, r.q[3].y, r.q[3].z, r.q[3].w};
  G1Aff29 p;
  p.x = unpack29<Q29>(x);
  p.y = unpack29<Q29>(y);
  return r.neg ? aff29_neg(p) : p;
}

// One thread per chunk of T consecutive sorted entries of a physical window
// past its z zero digits (the last chunk may be shorter), so every lane of a
// wave adds about the same number of points whatever the bucket sizes and
// the share of zero scalars.  The chunk's run inside bucket b is task
// k = chunk - (bstart[b] - z) / T of that bucket; its partial sum, mixed
// XYZZ + affine additions, goes to the task's slot.  The next entry's point and key and the index
// after it are loaded before each addition, so their latency hides behind it.
__device__ __forceinline__ void tasks_body(const MsmJob& J, u64 gi) {
  const MsmShape& s = J.s;
  const u64 cpw = chunks_per_window(s);
  const u32 pw = (u32)(gi / cpw);
  if (pw >= s.n_windows) return;
  const u64 c = gi - (u64)pw * cpw;
  const u32 z = gld(J.bend + (u64)pw * (s.nb + 1));
  if (z + c * s.T >= s.seg_len) return;
  const u32 lo = (u32)(z + c * s.T);
  const u32 hi = (u32)(lo + s.T < s.seg_len ? lo + s.T : s.seg_len);
  const u32* kk = sorted_keys(J) + (u64)pw * s.seg_len;
  const u32* vv = sorted_vals(J) + (u64)pw * s.seg_len;
  const u32* to = J.toff + (u64)pw * (s.nb + 2);
  const u32* bs = J.bstart + (u64)pw * (s.nb + 1);
  const u64 pbase = (u64)pw * s.task_cap;
  G1Xyzz29 acc = G1Xyzz29::identity();
  u32 b = gld(kk + lo);
  RawEntry cur = load_entry(J, gld(vv + lo));
  u32 key1 = lo + 1 < hi ? gld(kk + lo + 1) : 0u;
  u32 val1 = lo + 1 < hi ? gld(vv + lo + 1) : 0u;
  for (u32 j = lo; j < hi; j++) {
    RawEntry nxt = cur;
    u32 bn = b;
    if (j + 1 < hi) {
      nxt = load_entry(J, val1);
      bn = key1;
    }
    if (j + 2 < hi) {
      key1 = gld(kk + j + 2);
      val1 = gld(vv + j + 2);
    }
    if (b != 0) acc = xyzz29_madd(acc, unpack_entry(cur));
    if (j + 1 == hi || bn != b) {
      if (b != 0)
        store_xyzz29_g(J.partA + kPt * (pbase + to[b] + (u32)(c - (bs[b] - z) / s.T)), acc);
      acc = G1Xyzz29::identity();
    }
    cur = nxt;
    b = bn;
  }
}

// level l of the per-bucket partial-sum tree: compact index o of window pw
// (o < loff[nb+1]) is group k = o - loff[b] of bucket b; it sums the live
// entries src[toff_b + kG .. + G) of the previous level into dst[toff_b + k].
// Threads past the window's live groups exit at once.
__host__ __device__ inline u64 combine_cap(const MsmShape& s, u32 level) {
  return (s.task_cap >> (s.g * level)) + s.nb + 1;
}

__device__ __forceinline__ void combine_body(const MsmJob& J, u32 level, u64 gi) {
  const MsmShape& s = J.s;
  if (level > s.levels) return;
  const u64 cap = combine_cap(s, level);
  const u32 pw = (u32)(gi / cap);
  if (pw >= s.n_windows || level_idle(J, level, pw)) return;
  const u32 o = (u32)(gi - (u64)pw * cap);
  const u32* lo = J.loff + (u64)pw * (s.nb + 2);
  if (o >= lo[s.nb + 1]) return;
  const u32 b = find_bucket(lo, s.nb, o);
  const u32 k = o - lo[b];
  const u32* to = J.toff + (u64)pw * (s.nb + 2);
  const u32 nt = to[b + 1] - to[b];
  const u32 shp = s.g * (level - 1);
  const u32 n_prev = (u32)(((u64)nt + (1ull << shp) - 1) >> shp);
  const u32* src = (level & 1) ? J.partA : J.partB;
  u32* dst = (level & 1) ? J.partB : J.partA;
  const u64 base = (u64)pw * s.task_cap + to[b];
  const u32 G = 1u << s.g;
  G1Xyzz29 acc = load_xyzz29_g(src + kPt * (base + (u64)k * G));
  const u32 jn = n_prev - k * G < G ? n_prev - k * G : G;
  for (u32 j = 1; j < jn; j++)
    acc = xyzz29_add(acc, load_xyzz29_g(src + kPt * (base + (u64)k * G + j)));
  store_xyzz29_g(dst + kPt * (base + k), acc);
}

// Bucket b (1..nb) of window pw: this part's total (identity for an empty
// bucket) stored into (bmode 1) or added onto (bmode 2) the kept sums.
__device__ __forceinline__ void accum_body(const MsmJob& J, u64 i) {
  const MsmShape& s = J.s;
  if (J.bacc == nullptr || i >= (u64)s.n_windows * s.nb) return;
  const u32 pw = (u32)(i / s.nb), b = (u32)(i % s.nb) + 1;
  const u32* to = J.toff + (u64)pw * (s.nb + 2);
  const u32 nt = to[b + 1] - to[b];
  u32* dst = J.bacc + kPt * i;
  G1Xyzz29 v = G1Xyzz29::identity();
  if (nt) {
    const u32* pp = (bucket_levels(nt, s.g) & 1) ? J.partB : J.partA;
    v = load_xyzz29_g(pp + kPt * ((u64)pw * s.task_cap + to[b]));
  }
  if (J.bmode == 1) {
    store_xyzz29_g(dst, v);
  } else if (nt) {
    store_xyzz29_g(dst, xyzz29_add(load_xyzz29_g(dst), v));
  }
}

// Window sum S = sum_b b * B_b (b = 1 .. nb) in three shallow stages.
// Segment s holds buckets m s + 1 .. m s + m (m = seg = 2^sbits):
//   T_s = sum_k (k + 1) B_(ms+k+1),  R_s = sum_k B_(ms+k+1),
//   S = sum_s T_s + m sum_s s R_s = sum_s T_s + m sum_l 2^l V_l,
//   V_l = sum of R_s over the s with bit l set (l < lbits = log2 nseg).
// reduce: one thread per segment (running sums over m buckets);
// sum:    lbits + 1 sums (the T_s, then each V_l), one or two tree levels;
// final:  partials -> a_0 = sum T, a_1 .. a_(sbits-1) = 0, a_(sbits+l) = V_l,
//         S = sum_i 2^i a_i (each a_i doubled by its own lane, then a
//         tree of additions), then Horner over physical windows and affine.
__device__ __forceinline__ void reduce_body(const MsmJob& J, u64 t) {
  const MsmShape& s = J.s;
  if (t >= (u64)s.n_windows * s.nseg) return;
  const u32 pw = (u32)(t / s.nseg), sg = (u32)(t % s.nseg);
  const u32 lo = sg * s.seg;
  const u32* to = J.toff + (u64)pw * (s.nb + 2);
  const u64 pbase = (u64)pw * s.task_cap;
  G1Xyzz29 run = G1Xyzz29::identity(), tot = G1Xyzz29::identity();
  for (int k = (int)s.seg - 1; k >= 0; k--) {
    const u32 b = lo + k + 1;
    if (J.bacc) {
      run = xyzz29_add(run, load_xyzz29_g(J.bacc + kPt * ((u64)pw * s.nb + b - 1)));
    } else {
      const u32 nt = to[b + 1] - to[b];
      if (nt) {
        const u32* pp = (bucket_levels(nt, s.g) & 1) ? J.partB : J.partA;
        run = xyzz29_add(run, load_xyzz29_g(pp + kPt * (pbase + to[b])));
      }
    }
    tot = xyzz29_add(tot, run);
  }
  store_xyzz29_g(J.rparts + kPt * ((u64)(2 * pw) * s.nseg + sg), tot);
  store_xyzz29_g(J.rparts + kPt * ((u64)(2 * pw + 1) * s.nseg + sg), run);
}

// pairwise tree over the 256 XYZZ values red[0..255]: after it red[0] is
// their sum (blockDim.x == 256).  The first level adds on one lane per pair;
// from 64 pairs down each addition runs on a group of four lanes
// (xyzz29_add_quad), whose lane 0 writes the slot after a barrier.
__device__ __forceinline__ void block_tree_sum(u32* red) {
  const u32 t = threadIdx.x;
  if (t < 128)
    store_xyzz29(red + kPt * t,
                 xyzz29_add(load_xyzz29(red + kPt * t), load_xyzz29(red + kPt * (t + 128))));
  __syncthreads();
  const u32 q = t & 3, qt = t >> 2;
  for (u32 d = 64; d > 0; d >>= 1) {
    const bool live = qt < d;
    G1Xyzz29 r;
    if (live) r = xyzz29_add_quad(load_xyzz29(red + kPt * qt), load_xyzz29(red + kPt * (qt + d)), q);
    __syncthreads();
    if (live && q == 0) store_xyzz29(red + kPt * qt, r);
    __syncthreads();
  }
}

// Level-0 work-groups (and live partial slots) of window sum q: g0 for the
// T_s (q = 0), half as many for a V_l, which has half the inputs.
__host__ __device__ __forceinline__ u32 sum_groups(const MsmShape& s, u32 q) {
  return q == 0 || s.g0 == 1 ? s.g0 : s.g0 / 2;
}

// Level 0, work-group lb: window pw, sum q (0: the T_s; q = l + 1: V_l),
// inputs [1024 g, 1024 g + 1024) of that sum, four per thread (two pairs,
// then their sum), then a tree; the work-groups of a window are the g0 of
// the T_s, then sum_groups(s, 1) for each V_l.  Level 1, work-group lb:
// window pw, sum q, its level-0 partials.
__device__ __forceinline__ void sum_body(const MsmJob& J, u32 level, u64 lb, u32* red) {
  const MsmShape& s = J.s;
  const u32 t = threadIdx.x, nq = s.lbits + 1;
  G1Xyzz29 acc = G1Xyzz29::identity();
  __syncthreads();
  if (level == 0) {
    const u32 gh = sum_groups(s, 1), per_win = s.g0 + s.lbits * gh;
    const u32 pw = (u32)(lb / per_win), r = (u32)(lb % per_win);
    const u32 q = r < s.g0 ? 0 : 1 + (r - s.g0) / gh;
    const u32 g = r < s.g0 ? r : (r - s.g0) % gh;
    const u32 count = q == 0 ? s.nseg : s.nseg / 2;
    const u32* src = J.rparts + kPt * ((u64)(2 * pw + (q ? 1 : 0)) * s.nseg);
    const u32 l = q ? q - 1 : 0;
    // both inputs of a pair loaded before its addition
    G1Xyzz29 pair[2];
#pragma unroll
    for (u32 hp = 0; hp < 2; hp++) {
      G1Xyzz29 in[2];
#pragma unroll
      for (u32 h = 0; h < 2; h++) {
        const u32 i = g * 1024 + (2 * hp + h) * 256 + t;
        in[h] = G1Xyzz29::identity();
        if (i < count) {
          // q > 0: the i-th segment index with bit l set
          const u32 sg = q == 0 ? i : (((i >> l) << (l + 1)) | (1u << l) | (i & ((1u << l) - 1)));
          in[h] = load_xyzz29_g(src + kPt * (u64)sg);
        }
      }
      pair[hp] = xyzz29_add(in[0], in[1]);
    }
    acc = xyzz29_add(pair[0], pair[1]);
    store_xyzz29(red + kPt * t, acc);
    __syncthreads();
    block_tree_sum(red);
    if (t < kPt) J.sparts[kPt * (((u64)pw * nq + q) * s.g0 + g) + t] = red[t];
  } else {
    const u64 v = lb;  // pw * nq + q
    const u32 ng = sum_groups(s, (u32)(v % nq));
    for (u32 i = t; i < ng; i += 256)
      acc = xyzz29_add(acc, load_xyzz29_g(J.sparts + kPt * (v * s.g0 + i)));
    store_xyzz29(red + kPt * t, acc);
    __syncthreads();
    block_tree_sum(red);
    if (t < kPt) J.sparts1[kPt * v + t] = red[t];
  }
}

// one work-group per job (blockDim.x == 256); the shape guarantees
// n_windows * (lbits + sbits) * fin_g <= 256.  Everything here is a chain of
// dependent EC operations on a few lanes, where one wave's latency counts,
// not the chip's product rate: each value is worked on by a group of four
// lanes (xyzz29_dbl_quad / xyzz29_add_quad: three or four rounds of
// products instead of nine or fourteen in a row) wherever the work-group
// has four lanes per value, otherwise by one lane; Horner over the physical
// windows always runs on lanes 0..3.  The group's lane 0 writes a value's
// slot, after a barrier (every lane has read its operands by then).
__device__ __forceinline__ G1Xyzz29 final_add(const G1Xyzz29& a, const G1Xyzz29& b, bool quad,
                                              u32 q) {
  return quad ? xyzz29_add_quad(a, b, q) : xyzz29_add(a, b);
}
__device__ __forceinline__ void final_body(const MsmJob& J, u32* red) {
  const MsmShape& s = J.s;
  const u32 t = threadIdx.x;
  __syncthreads();
  if (s.N == 0) {
    if (t < 16) J.result[t] = 0;
    return;
  }
  const u32 nv = s.lbits + s.sbits, G = s.fin_g, nval = s.n_windows * nv;
  const u32* parts = s.sum_levels == 2 ? J.sparts1 : J.sparts;
  // value (pw, i) at red[(pw * nv + i) * G], its sum partials in the next G - 1 slots
  {
    const bool quad = nval * G * 4 <= 256;
    const u32 q = quad ? t & 3 : 0, vt = quad ? t >> 2 : t;
    if (vt < nval * G && q == 0) {
      const u32 gi = vt % G, i = (vt / G) % nv, pw = vt / (G * nv);
      G1Xyzz29 v = G1Xyzz29::identity();
      const u32 qs = i == 0 ? 0 : i - s.sbits + 1;
      // with one sum level, slots past the sum's live partials are unused
      if ((i == 0 || i >= s.sbits) && (G == 1 || gi < sum_groups(s, qs)))
        v = load_xyzz29_g(parts + kPt * (((u64)pw * (s.lbits + 1) + qs) * G + gi));
      store_xyzz29(red + kPt * vt, v);
    }
    __syncthreads();
    for (u32 d = 1; d < G; d <<= 1) {
      const bool live = vt < nval * G && (vt % (2 * d)) == 0;
      G1Xyzz29 r;
      if (live)
        r = final_add(load_xyzz29(red + kPt * vt), load_xyzz29(red + kPt * (vt + d)), quad, q);
      __syncthreads();
      if (live && q == 0) store_xyzz29(red + kPt * vt, r);
      __syncthreads();
    }
  }
  // S = sum_i 2^i a_i per window: value (pw, i) is doubled i times (the
  // values side by side, so the chain is nv - 1 doublings deep), then a
  // tree adds the nv terms
  {
    const bool quad = nval * 4 <= 256;
    const u32 q = quad ? t & 3 : 0, vt = quad ? t >> 2 : t;
    // a_1 .. a_(sbits-1) are the identity
    const bool dlive = vt < nval && vt % nv >= s.sbits;
    G1Xyzz29 b;
    if (dlive) {
      b = load_xyzz29(red + kPt * (vt * G));
      for (u32 k = 0, i = vt % nv; k < i; k++) b = quad ? xyzz29_dbl_quad(b, q) : xyzz29_dbl(b);
    }
    __syncthreads();
    if (dlive && q == 0) store_xyzz29(red + kPt * (vt * G), b);
    __syncthreads();
    for (u32 h = 1; h < nv; h <<= 1) {
      const bool live = vt < nval && ((vt % nv) % (2 * h)) == 0 && vt % nv + h < nv;
      G1Xyzz29 r;
      if (live)
        r = final_add(load_xyzz29(red + kPt * (vt * G)), load_xyzz29(red + kPt * ((vt + h) * G)),
                      quad, q);
      __syncthreads();
      if (live && q == 0) store_xyzz29(red + kPt * (vt * G), r);
      __syncthreads();
    }
  }
  if (t >= 4) return;
  // Horner over the physical windows on lanes 0..3: (n_windows - 1) c
  // dependent doublings
  G1Xyzz29 acc = load_xyzz29(red + kPt * ((s.n_windows - 1) * nv * G));
  for (int pw = (int)s.n_windows - 2; pw >= 0; pw--) {
    for (u32 k = 0; k < s.c; k++) acc = xyzz29_dbl_quad(acc, t);
    acc = xyzz29_add_quad(acc, load_xyzz29(red + kPt * ((u32)pw * nv * G)), t);
  }
  if (t == 0) store_aff(J.result, xyzz_to_aff(xyzz_from_29(acc)));
}

// ------------------------------------------------------- batched kernels
// Every stage is one launch over the concatenated block space of all jobs;
// prefix[j] = first block of job j (jobs without work contribute none), and
// work-groups stride over that space.
__device__ __forceinline__ u32 find_job(const u32* prefix, u32 njobs, u32 t) {
  u32 lo = 0, hi = njobs - 1;
  while (lo < hi) {
    u32 mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

enum MsmStage : u32 {
  MS_DIGITS = 0, MS_HIST, MS_SCAN_A, MS_SCAN_B, MS_SCAN_C, MS_SCATTER, MS_ZERO, MS_BOUNDS,
  MS_TASKS, MS_COMBINE, MS_REDUCE, MS_SUM, MS_FINAL, MS_CONVERT, MS_ACCUM
};
// scan stage argument: ScanKind | (pass or level) << 8
static inline u32 scan_arg(u32 kind, u32 x) { return kind | (x << 8); }

#define IMP_BATCH_LOOP(BODY)                                        \
  const u32 total = prefix[njobs];                                  \
  for (u32 t = blockIdx.x; t < total; t += gridDim.x) {             \
    u32 j = find_job(prefix, njobs, t);                             \
    const MsmJob& J = jobs[j];                                      \
    const u64 lb = t - prefix[j];                                   \
    BODY;                                                           \
  }

// Work-groups are dealt to the 8 XCDs round-robin (block b runs on XCD
// b mod 8).  xcd_order(b, n) renumbers the blocks of a one-block-per-item
// grid so that every XCD takes one contiguous range of [0, n): consecutive
// sort tiles then share one L2, where the adjacent partial-line runs they
// write for one digit merge before going to HBM.  A bijection on [0, n).
__device__ __forceinline__ u32 xcd_order(u32 b, u32 n) {
  const u32 x = b & 7, k = b >> 3, q = n >> 3, r = n & 7;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

#define IMP_BATCH_LOOP_XCD(BODY)                                    \
  const u32 total = prefix[njobs];                                  \
  const bool remap = gridDim.x >= total;                            \
  for (u32 t0 = blockIdx.x; t0 < total; t0 += gridDim.x) {          \
    const u32 t = remap ? xcd_order(t0, total) : t0;                \
    u32 j = find_job(prefix, njobs, t);                             \
    const MsmJob& J = jobs[j];                                      \
    const u64 lb = t - prefix[j];                                   \
    BODY;                                                           \
  }

__global__ __launch_bounds__(256) void kb_convert(const MsmJob* jobs, u32 njobs, const u32* prefix) {
  IMP_BATCH_LOOP(convert_body(J, lb * blockDim.x + threadIdx.x))
}
__global__ __launch_bounds__(256) void kb_digits(const MsmJob* jobs, u32 njobs, const u32* prefix) {
  IMP_BATCH_LOOP(digits_body(J, 4 * (lb * blockDim.x + threadIdx.x)))
}
__global__ __launch_bounds__(256) void kb_hist(const MsmJob* jobs, u32 njobs, const u32* prefix,
                                               u32 pass) {
  __shared__ u32 cnt[4 * kRadix];
  IMP_BATCH_LOOP_XCD(hist_body(J, pass, lb, cnt))
}
__global__ __launch_bounds__(256) void kb_scan_a(const MsmJob* jobs, u32 njobs,
                                                 const u32* prefix, u32 arg) {
  __shared__ u32 part[256];
  IMP_BATCH_LOOP(scan_a_body(J, arg & 255, arg >> 8, lb, part))
}
__global__ __launch_bounds__(256) void kb_scan_b(const MsmJob* jobs, u32 njobs,
                                                 const u32* prefix, u32 arg) {
  __shared__ u32 part[256];
  IMP_BATCH_LOOP(scan_b_body(J, arg & 255, arg >> 8, lb, part))
}
__global__ __launch_bounds__(256) void kb_scan_c(const MsmJob* jobs, u32 njobs,
                                                 const u32* prefix, u32 arg) {
  __shared__ u32 part[256];
  IMP_BATCH_LOOP(scan_c_body(J, arg & 255, arg >> 8, lb, part))
}
__global__ __launch_bounds__(256) void kb_scatter(const MsmJob* jobs, u32 njobs,
                                                  const u32* prefix, u32 pass) {
  __shared__ ScatterLds lds;
  IMP_BATCH_LOOP_XCD(scatter_body(J, pass, lb, lds))
}
__global__ __launch_bounds__(256) void kb_zero(const MsmJob* jobs, u32 njobs, const u32* prefix) {
  IMP_BATCH_LOOP(zero_body(J, lb * blockDim.x + threadIdx.x))
}
__global__ __launch_bounds__(256) void kb_bounds(const MsmJob* jobs, u32 njobs, const u32* prefix) {
  IMP_BATCH_LOOP(bounds_body(J, lb))
}
__global__ __launch_bounds__(256) void kb_tasks(const MsmJob* jobs, u32 njobs, const u32* prefix) {
  IMP_BATCH_LOOP(tasks_body(J, lb * blockDim.x + threadIdx.x))
}
__global__ __launch_bounds__(256) void kb_combine(const MsmJob* jobs, u32 njobs,
                                                  const u32* prefix, u32 level) {
  IMP_BATCH_LOOP(combine_body(J, level, lb * blockDim.x + threadIdx.x))
}
__global__ __launch_bounds__(256) void kb_accum(const MsmJob* jobs, u32 njobs, const u32* prefix) {
  IMP_BATCH_LOOP(accum_body(J, lb * blockDim.x + threadIdx.x))
}
__global__ __launch_bounds__(256) void kb_reduce(const MsmJob* jobs, u32 njobs, const u32* prefix) {
  IMP_BATCH_LOOP(reduce_body(J, lb * blockDim.x + threadIdx.x))
}
__global__ __launch_bounds__(256) void kb_sum(const MsmJob* jobs, u32 njobs, const u32* prefix,
                                              u32 level) {
  __shared__ u32 red[256 * kPt];
  IMP_BATCH_LOOP(sum_body(J, level, lb, red))
}
__global__ __launch_bounds__(256) void kb_final(const MsmJob* jobs, u32 njobs, const u32* prefix) {
  __shared__ __align__(16) u32 red[256 * kPt];
  IMP_BATCH_LOOP(final_body(J, red))
}

// table t = 2^(t * step) * table 0 (affine), t = 1 .. n_pre-1
__global__ void k_msm_precompute(u32* tables, u64 N, u32 n_pre, u32 step) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  G1Aff p = load_aff(tables + 16 * i);
  G1Xyzz acc = aff_to_xyzz(p);
  for (u32 t = 1; t < n_pre; t++) {
    for (u32 k = 0; k < step; k++) acc = xyzz_dbl(acc);
    store_aff(tables + 16 * ((u64)t * N + i), xyzz_to_aff(acc, true));
  }
}

// affine R-form points -> table format (in place allowed: each point is read
// once by the thread that writes it)
__global__ void k_msm_to_table(u32* dst, const u32* src, u64 count) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  aff_to_table(dst + 16 * i, load_aff(src + 16 * i));
}

static hipError_t to_table(u32* dst, const u32* src, u64 count, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_msm_to_table, dim3((u32)((count + 255) / 256)), dim3(256), 0, stream, dst,
                     src, count);
  return hipGetLastError();
}

// ------------------------------------------------------------------ plan
// A launch list over a set of jobs: stage kind, argument (pass or level) and
// per-job block counts.
struct StageLaunch {
  u32 kind, arg, total;
};

static inline u64 cdiv(u64 a, u64 b) { return (a + b - 1) / b; }

static u32 stage_blocks(const MsmShape& s, u32 kind, u32 arg) {
  if (s.N == 0) return kind == MS_FINAL ? 1u : 0u;
  switch (kind) {
    case MS_CONVERT: return s.n_pre == 1 ? (u32)cdiv(s.N, 256) : 0u;
    case MS_DIGITS: return (u32)cdiv(s.N, 1024);  // four points per thread
    case MS_HIST:
    case MS_SCATTER: return arg < s.passes ? s.tiles * s.n_windows : 0;
    case MS_SCAN_A:
    case MS_SCAN_B:
    case MS_SCAN_C: {
      const u32 sk = arg & 255, x = arg >> 8;
      const bool live = sk == SC_HIST ? x < s.passes : (sk == SC_LEVEL ? x <= s.levels : true);
      if (!live) return 0;
      return kind == MS_SCAN_B ? s.n_windows : s.n_windows * scan_chunks(s, sk);
    }
    case MS_ZERO: return (u32)cdiv(2ull * s.n_windows * (s.nb + 1) + s.n_windows, 256);
    case MS_BOUNDS: return (u32)(s.n_windows * bounds_blocks_per_window(s));
    case MS_TASKS: return (u32)cdiv((u64)s.n_windows * chunks_per_window(s), 256);
    case MS_COMBINE:
      return arg <= s.levels ? (u32)cdiv((u64)s.n_windows * combine_cap(s, arg), 256) : 0;
    case MS_REDUCE: return (u32)cdiv((u64)s.n_windows * s.nseg, 256);
    case MS_ACCUM: return (u32)cdiv((u64)s.n_windows * s.nb, 256);
    case MS_SUM:
      if (arg == 0) return s.n_windows * (s.g0 + s.lbits * sum_groups(s, 1));
      return s.sum_levels == 2 ? s.n_windows * (s.lbits + 1) : 0u;
    default: return 1;
  }
}

// part: 0 a whole MSM; 1 a part that only accumulates its bucket totals;
// 2 the last part: accumulates, then reduces the kept sums to the result.
// The stages before *points_at read only the scalars (digits, the sort,
// the bucket bounds and task offsets); the base conversion and everything
// after it read the points, so a caller streaming the points in can launch
// the first stages while the points are still being copied.
static void plan_stages(const std::vector<MsmShape>& shapes, std::vector<StageLaunch>& out,
                        std::vector<u32>& prefix, u32 part = 0, u32* points_at = nullptr) {
  u32 maxp = 0, maxl = 0;
  for (auto& s : shapes) {
    if (s.passes > maxp) maxp = s.passes;
    if (s.levels > maxl) maxl = s.levels;
  }
  std::vector<std::pair<u32, u32>> kinds;
  kinds.push_back({MS_DIGITS, 0});
  auto scan3 = [&](u32 sk, u32 x) {
    kinds.push_back({MS_SCAN_A, scan_arg(sk, x)});
    kinds.push_back({MS_SCAN_B, scan_arg(sk, x)});
    kinds.push_back({MS_SCAN_C, scan_arg(sk, x)});
  };
  for (u32 p = 0; p < maxp; p++) {
    kinds.push_back({MS_HIST, p});
    scan3(SC_HIST, p);
    kinds.push_back({MS_SCATTER, p});
  }
  kinds.push_back({MS_ZERO, 0});
  kinds.push_back({MS_BOUNDS, 0});
  scan3(SC_TASKS, 0);
  if (points_at) *points_at = (u32)kinds.size();
  kinds.push_back({MS_CONVERT, 0});
  kinds.push_back({MS_TASKS, 0});
  for (u32 l = 1; l <= maxl; l++) {
    scan3(SC_LEVEL, l);
    kinds.push_back({MS_COMBINE, l});
  }
  if (part) kinds.push_back({MS_ACCUM, 0});
  if (part != 1) {
    kinds.push_back({MS_REDUCE, 0});
    kinds.push_back({MS_SUM, 0});
    kinds.push_back({MS_SUM, 1});
    kinds.push_back({MS_FINAL, 0});
  }
  const u32 n = (u32)shapes.size();
  out.clear();
  prefix.assign(kinds.size() * (n + 1), 0);
  for (size_t k = 0; k < kinds.size(); k++) {
    u32 run = 0;
    for (u32 j = 0; j < n; j++) {
      prefix[k * (n + 1) + j] = run;
      run += stage_blocks(shapes[j], kinds[k].first, kinds[k].second);
    }
    prefix[k * (n + 1) + n] = run;
    out.push_back(StageLaunch{kinds[k].first, kinds[k].second, run});
  }
}

// timer of each stage kind; a scan is booked with the stage it serves (the
// radix histograms with the sort, the task offsets with the bucket bounds,
// the combine levels with the combine)
static const u32 kTimerOf[] = {K_MSM_DIGITS,  K_MSM_SORT,   K_MSM_SORT, K_MSM_SORT, K_MSM_SORT,
                               K_MSM_SORT,    K_MSM_SCAN,   K_MSM_SCAN, K_MSM_BUCKETS,
                               K_MSM_COMBINE, K_MSM_REDUCE, K_MSM_SUM,  K_MSM_FINAL,
                               K_MSM_DIGITS,  K_MSM_COMBINE};
static u32 stage_timer(const StageLaunch& st) {
  if (st.kind == MS_SCAN_A || st.kind == MS_SCAN_B || st.kind == MS_SCAN_C) {
    const u32 sk = st.arg & 255;
    return sk == SC_HIST ? K_MSM_SORT : (sk == SC_TASKS ? K_MSM_SCAN : K_MSM_COMBINE);
  }
  return kTimerOf[st.kind];
}

static hipError_t launch_stages(const MsmJob* jobs, u32 njobs, const u32* prefix_dev,
                                const StageLaunch* st, u32 nst, u32 max_blocks,
                                hipStream_t stream) {
  for (u32 k = 0; k < nst; k++) {
    u32 total = st[k].total;
    if (total == 0) continue;
    u32 grid = max_blocks && total > max_blocks ? max_blocks : total;
    const u32* pf = prefix_dev + (size_t)k * (njobs + 1);
    TimedScope ts(stage_timer(st[k]), stream);
    switch (st[k].kind) {
      case MS_CONVERT: hipLaunchKernelGGL(kb_convert, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf); break;
      case MS_ACCUM: hipLaunchKernelGGL(kb_accum, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf); break;
      case MS_DIGITS: hipLaunchKernelGGL(kb_digits, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf); break;
      case MS_HIST: hipLaunchKernelGGL(kb_hist, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf, st[k].arg); break;
      case MS_SCAN_A: hipLaunchKernelGGL(kb_scan_a, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf, st[k].arg); break;
      case MS_SCAN_B: hipLaunchKernelGGL(kb_scan_b, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf, st[k].arg); break;
      case MS_SCAN_C: hipLaunchKernelGGL(kb_scan_c, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf, st[k].arg); break;
      case MS_SCATTER: hipLaunchKernelGGL(kb_scatter, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf, st[k].arg); break;
      case MS_ZERO: hipLaunchKernelGGL(kb_zero, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf); break;
      case MS_BOUNDS: hipLaunchKernelGGL(kb_bounds, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf); break;
      case MS_TASKS: hipLaunchKernelGGL(kb_tasks, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf); break;
      case MS_COMBINE: hipLaunchKernelGGL(kb_combine, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf, st[k].arg); break;
      case MS_REDUCE: hipLaunchKernelGGL(kb_reduce, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf); break;
      case MS_SUM: hipLaunchKernelGGL(kb_sum, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf, st[k].arg); break;
      default: hipLaunchKernelGGL(kb_final, dim3(grid), dim3(256), 0, stream, jobs, njobs, pf); break;
    }
    IMP_CHECK(hipGetLastError());
  }
  return hipSuccess;
}

// descriptor area at the front of a workspace: jobs, then stage prefixes
static const size_t kMaxStages = 128;
static size_t desc_bytes(u32 njobs) {
  return align256(sizeof(MsmJob) * njobs) + align256(kMaxStages * (njobs + 1) * 4);
}

// A single MSM is some 45 launches, many of them a few microseconds long, so
// issuing them one by one leaves the GPU waiting on the host.  Their
// sequence depends only on the shape and on the workspace (every kernel
// reads its descriptor from the workspace, which k_msm_upload rewrites
// before each call), so it is captured once into a HIP graph per (device,
// workspace, stage list) and replayed.  Direct launches are kept while the
// kernel timing registry is on (its events would be captured) and when the
// caller's stream is itself being captured.  At most kGraphCache graphs are
// kept; the oldest is destroyed after a device synchronisation.  The cache
// is never torn down (graphs destroyed after the runtime's own teardown at
// exit would be unsafe).
static const size_t kGraphCache = 32;

struct StageGraphCache {
  std::mutex mu;
  std::vector<std::pair<std::string, hipGraphExec_t>> entries;  // oldest first
};

static StageGraphCache& stage_graphs() {
  static StageGraphCache* c = new StageGraphCache();
  return *c;
}

static hipError_t capture_stages(const MsmJob* jobs_d, const u32* prefix_d,
                                 const std::vector<StageLaunch>& st, hipGraphExec_t* out) {
  hipStream_t cap = nullptr;
  IMP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  hipError_t le = hipSuccess;
  if (e == hipSuccess) {
    le = launch_stages(jobs_d, 1, prefix_d, st.data(), (u32)st.size(), 0, cap);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cap, &g);
    if (e == hipSuccess && le == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
  }
  (void)hipStreamDestroy(cap);
  return e != hipSuccess ? e : le;
}

static hipError_t launch_cached(const MsmJob* jobs_d, const u32* prefix_d,
                                const std::vector<StageLaunch>& st, hipStream_t stream) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (timing_on() || hipStreamIsCapturing(stream, &cs) != hipSuccess ||
      cs != hipStreamCaptureStatusNone)
    return launch_stages(jobs_d, 1, prefix_d, st.data(), (u32)st.size(), 0, stream);
  int dev = 0;
  IMP_CHECK(hipGetDevice(&dev));
  std::string key((const char*)&dev, sizeof(dev));
  key.append((const char*)&jobs_d, sizeof(jobs_d));
  key.append((const char*)st.data(), st.size() * sizeof(StageLaunch));
  StageGraphCache& c = stage_graphs();
  std::lock_guard<std::mutex> lk(c.mu);
  for (size_t i = 0; i < c.entries.size(); i++) {
    if (c.entries[i].first == key) {
      auto hit = c.entries[i];
      c.entries.erase(c.entries.begin() + i);
      c.entries.push_back(hit);
      return hipGraphLaunch(hit.second, stream);
    }
  }
  hipGraphExec_t exec = nullptr;
  if (capture_stages(jobs_d, prefix_d, st, &exec) != hipSuccess || exec == nullptr) {
    (void)hipGetLastError();
    return launch_stages(jobs_d, 1, prefix_d, st.data(), (u32)st.size(), 0, stream);
  }
  if (c.entries.size() >= kGraphCache) {
    IMP_CHECK(hipDeviceSynchronize());
    (void)hipGraphExecDestroy(c.entries.front().second);
    c.entries.erase(c.entries.begin());
  }
  c.entries.push_back({key, exec});
  return hipGraphLaunch(exec, stream);
}

// A single MSM's descriptor and stage prefixes travel as kernel arguments to
// one small kernel that writes them into the workspace, in stream order: no
// host staging to keep alive, no synchronisation, and the call stays a pure
// sequence of launches (graph-capturable).
struct PrefixBlock {
  u32 v[kMaxStages * 2];
};

__global__ __launch_bounds__(256) void k_msm_upload(MsmJob J, PrefixBlock P, u32 np,
                                                   MsmJob* jd, u32* pd) {
  const u32* src = reinterpret_cast<const u32*>(&J);
  u32* dst = reinterpret_cast<u32*>(jd);
  for (u32 i = threadIdx.x; i < sizeof(MsmJob) / 4; i += blockDim.x) dst[i] = src[i];
  for (u32 i = threadIdx.x; i < np; i += blockDim.x) pd[i] = P.v[i];
}

static hipError_t msm_run(const u32* bases, u64 tstride, const u32* scalars, const MsmShape& s,
                          u32* result, char* ws, hipStream_t stream) {
  char* body = ws + desc_bytes(1);
  MsmJob J = make_msm_job(s, tstride, bases, scalars, result, body);
  std::vector<MsmShape> shapes(1, s);
  std::vector<StageLaunch> st;
  std::vector<u32> prefix;
  plan_stages(shapes, st, prefix);
  IMP_REQUIRE(prefix.size() <= kMaxStages * 2, "stage plan");
  MsmJob* jobs_d = (MsmJob*)ws;
  u32* prefix_d = (u32*)(ws + align256(sizeof(MsmJob)));
  PrefixBlock P;
  memset(&P, 0, sizeof(P));
  memcpy(P.v, prefix.data(), prefix.size() * 4);
  hipLaunchKernelGGL(k_msm_upload, dim3(1), dim3(256), 0, stream, J, P, (u32)prefix.size(),
                     jobs_d, prefix_d);
  IMP_CHECK(hipGetLastError());
  return launch_cached(jobs_d, prefix_d, st, stream);
}

static size_t msm_ws_total(const MsmShape& s) { return desc_bytes(1) + msm_layout(s).total; }

// c in [4, 20]; the final work-group holds lbits + sbits = c - 1 values for each of
// up to W = 254 / c + 1 windows, at most 255 over that range
static bool valid_window(u32 c) { return c >= 4 && c <= 20 && (254 / c + 1) * (c - 1) <= 256; }

static hipError_t precompute_device(u32* tables, u64 N, const MsmShape& s, hipStream_t stream) {
  if (s.n_pre <= 1 || N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_msm_precompute, dim3((u32)cdiv(N, 128)), dim3(128), 0, stream, tables, N,
                     s.n_pre, s.n_windows * s.c);
  return hipGetLastError();
}

// ---------------------------------------------------- host-memory contract
// The reference's msm (msm.cu:9-41, msm_impl.cuh:373-580,703-756): points
// and scalars in host memory, results back to host memory.  The points are
// split into `parts` ranges of part_len = ceil(len / parts) points; each
// part's tables go through one of `stage_points` device buffers and each
// (part, batch) scalar range through one of `stage_scalers` buffers, copied
// on a copy stream while the compute stream runs the MSM of the previous
// range (events order the reuse of every buffer).  Batches run
// `batch_per_run` at a time; each batch keeps one set of bucket sums across
// the parts: every part sorts its digits and sums its buckets, then adds the
// bucket totals onto the kept set, and only the last part of the run reduces
// the kept buckets to the batch's result (the reduction tail runs once per
// batch, not once per part).  Runs alternate the order of the parts, so the
// part resident at the end of one run is not copied again at the start of
// the next.  Everything runs on the first card (this build drives one GPU;
// the other cards' buffer sizes are 0).
struct HostPlan {
  MsmShape s;      // shape of a full part (part_len points)
  u64 len, part_len;
  u32 parts, bpr, sp, ss;
  size_t pts_stage, scl_stage, ws_body, op_desc, descs, bacc, finals, total;
  size_t off_pts, off_scl, off_ws, off_desc, off_bacc, off_fin;
};

static size_t op_desc_bytes() {
  return align256(sizeof(MsmJob)) + align256(kMaxStages * 2 * 4);
}

static HostPlan host_plan(u32 window, u32 target, u64 len, u32 bpr, u32 parts, u32 ss, u32 sp) {
  HostPlan h;
  h.len = len;
  h.parts = parts ? parts : 1;
  if ((u64)h.parts > len && len > 0) h.parts = (u32)len;
  h.part_len = len ? (len + h.parts - 1) / h.parts : 0;
  h.bpr = bpr ? bpr : 1;
  h.sp = sp ? sp : 1;
  if (h.sp > h.parts) h.sp = h.parts;
  h.ss = ss ? ss : 1;
  h.s = msm_shape(h.part_len, window, target);
  h.pts_stage = align256((size_t)h.s.n_pre * h.part_len * 64);
  h.scl_stage = align256((size_t)h.part_len * 32);
  h.ws_body = align256(msm_layout(h.s).total);
  h.op_desc = op_desc_bytes();
  h.descs = (size_t)h.parts * h.bpr * h.op_desc;
  // each batch of a run keeps its bucket sums across the parts
  h.bacc = h.parts > 1 ? align256((size_t)h.s.n_windows * h.s.nb * kPt * 4) : 0;
  h.finals = align256((size_t)h.bpr * 64);
  size_t o = 0;
  h.off_pts = o; o += h.sp * h.pts_stage;
  h.off_scl = o; o += h.ss * h.scl_stage;
  h.off_ws = o; o += h.ws_body;
  h.off_desc = o; o += h.descs;
  h.off_bacc = o; o += h.bpr * h.bacc;
  h.off_fin = o; o += h.finals;
  h.total = o;
  return h;
}

// Owns the streams and events of one call and puts the caller's device back
// on every exit path.
struct HostCallGuard {
  int prev = -1;
  hipStream_t st = nullptr, cp = nullptr;
  std::vector<hipEvent_t> evs;
  ~HostCallGuard() {
    if (st) (void)hipStreamSynchronize(st);
    if (cp) (void)hipStreamSynchronize(cp);
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    if (cp) (void)hipStreamDestroy(cp);
    if (st) (void)hipStreamDestroy(st);
    if (prev >= 0) (void)hipSetDevice(prev);
  }
  hipError_t event(hipEvent_t* e) {
    IMP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    evs.push_back(*e);
    return hipSuccess;
  }
};

static hipError_t msm_host(u32 window, u32 target, void* const* buffers,
                           unsigned long* buffer_sizes, unsigned long long len, u32 batch_per_run,
                           u32 parts, u32 stage_scalers, u32 stage_points,
                           unsigned int num_cards, const unsigned int* cards,
                           unsigned int const* const* h_points, unsigned int batches,
                           unsigned int const* const* h_scaler_batch,
                           unsigned int* const* h_result) {
  IMP_REQUIRE(valid_window(window), "window");
  IMP_REQUIRE(num_cards >= 1 && cards, "cards");
  const HostPlan h = host_plan(window, target, len, batch_per_run, parts, stage_scalers,
                               stage_points);
  IMP_REQUIRE((u64)h.s.n_pre * h.part_len < (1ull << 31), "point index range (n_pre * part)");
  if (buffers == nullptr) {
    IMP_REQUIRE(buffer_sizes, "buffer_sizes");
    for (u32 i = 0; i < num_cards; i++) buffer_sizes[i] = i == 0 ? h.total : 0;
    return hipSuccess;
  }
  IMP_REQUIRE(buffers[0] && h_points && (batches == 0 || (h_scaler_batch && h_result)),
              "null pointer");
  if (batches == 0) return hipSuccess;
  if (len == 0) {
    for (u32 b = 0; b < batches; b++) memset(h_result[b], 0, 64);
    return hipSuccess;
  }
  for (u32 t = 0; t < h.s.n_pre; t++) IMP_REQUIRE(h_points[t], "point table");
  HostCallGuard g;
  IMP_CHECK(hipGetDevice(&g.prev));
  IMP_CHECK(hipSetDevice((int)cards[0]));
  IMP_CHECK(hipStreamCreateWithFlags(&g.st, hipStreamNonBlocking));
  IMP_CHECK(hipStreamCreateWithFlags(&g.cp, hipStreamNonBlocking));
  char* buf = (char*)buffers[0];
  std::vector<hipEvent_t> pt_ready(h.sp), pt_free(h.sp), sc_ready(h.ss), sc_free(h.ss);
  for (u32 k = 0; k < h.sp; k++) {
    IMP_CHECK(g.event(&pt_ready[k]));
    IMP_CHECK(g.event(&pt_free[k]));
    IMP_CHECK(hipEventRecord(pt_free[k], g.st));
  }
  for (u32 k = 0; k < h.ss; k++) {
    IMP_CHECK(g.event(&sc_ready[k]));
    IMP_CHECK(g.event(&sc_free[k]));
    IMP_CHECK(hipEventRecord(sc_free[k], g.st));
  }
  std::vector<long long> resident(h.sp, -1);  // part held by each point stage
  u32 next_pt = 0, next_sc = 0;
  const u32 runs = (batches + h.bpr - 1) / h.bpr;
  std::vector<MsmJob> jobs;
  // stage lists by [the part's shape: full, or the shorter last range]
  // [what it does: a whole MSM (one part), accumulate only, accumulate and
  // finish] (the reference keeps its bucket sums across parts and reduces
  // once, msm_impl.cuh:373-580)
  std::vector<StageLaunch> stl[2][3];
  std::vector<u32> pfl[2][3];
  u32 pat[2][3];  // first stage that reads the points
  {
    std::vector<MsmShape> one(1, h.s);
    const u64 last_len = h.len - (u64)(h.parts - 1) * h.part_len;
    for (u32 sh = 0; sh < 2; sh++) {
      one[0] = sh ? msm_shape(last_len, window, target) : h.s;
      for (u32 m = 0; m < 3; m++) {
        plan_stages(one, stl[sh][m], pfl[sh][m], m, &pat[sh][m]);
        IMP_REQUIRE(pfl[sh][m].size() <= kMaxStages * 2, "stage plan");
      }
    }
  }
  for (u32 r = 0; r < runs; r++) {
    const u32 b0 = r * h.bpr, nb = batches - b0 < h.bpr ? batches - b0 : h.bpr;
    const bool fwd = (r & 1) == 0;
    // the run's operations: one MSM per (part, batch), descriptors uploaded
    // once (the previous run is complete: its results were read back)
    struct Op {
      u32 part, batch, pstage;
    };
    std::vector<Op> ops;
    std::vector<u32> part_stage(h.parts);
    {
      std::vector<long long> res = resident;
      u32 np = next_pt;
      for (u32 q = 0; q < h.parts; q++) {
        const u32 p = fwd ? q : h.parts - 1 - q;
        u32 k = h.sp;
        for (u32 i = 0; i < h.sp; i++)
          if (res[i] == (long long)p) k = i;
        if (k == h.sp) {
          k = np;
          np = (np + 1) % h.sp;
          res[k] = p;
        }
        part_stage[p] = k;
        for (u32 j = 0; j < nb; j++) ops.push_back(Op{p, j, k});
      }
    }
    jobs.assign(ops.size(), MsmJob());
    std::vector<u32> mode(ops.size(), 0);
    char* desc = buf + h.off_desc;
    const u32 first_part = ops.empty() ? 0 : ops.front().part;
    const u32 last_part = ops.empty() ? 0 : ops.back().part;
    for (size_t o = 0; o < ops.size(); o++) {
      const Op& op = ops[o];
      const u64 cur = op.part + 1 == h.parts ? h.len - (u64)op.part * h.part_len : h.part_len;
      const MsmShape s = msm_shape(cur, window, target);
      u32* result = (u32*)(buf + h.off_fin) + 16 * op.batch;
      // scalar stage pointers are patched below (the stage is picked when
      // the copy is issued, in the same order)
      jobs[o] = make_msm_job(s, h.part_len, (const u32*)(buf + h.off_pts + op.pstage * h.pts_stage),
                             nullptr, result, buf + h.off_ws);
      if (h.parts > 1) {
        jobs[o].bacc = (u32*)(buf + h.off_bacc + op.batch * h.bacc);
        jobs[o].bmode = op.part == first_part ? 1u : 2u;
        mode[o] = op.part == last_part ? 2u : 1u;
      }
    }
    {
      u32 ns = next_sc;
      for (size_t o = 0; o < ops.size(); o++) {
        jobs[o].scalars = (const u32*)(buf + h.off_scl + ns * h.scl_stage);
        ns = (ns + 1) % h.ss;
      }
    }
    for (size_t o = 0; o < ops.size(); o++) {
      char* d = desc + o * h.op_desc;
      IMP_CHECK(hipMemcpyAsync(d, &jobs[o], sizeof(MsmJob), hipMemcpyHostToDevice, g.st));
      const std::vector<u32>& pf = pfl[ops[o].part + 1 == h.parts ? 1 : 0][mode[o]];
      IMP_CHECK(hipMemcpyAsync(d + align256(sizeof(MsmJob)), pf.data(), pf.size() * 4,
                               hipMemcpyHostToDevice, g.st));
    }
    IMP_CHECK(hipStreamSynchronize(g.st));
    // issue copies and MSMs in operation order.  On the copy stream an
    // operation's scalars go first, then (for the first operation of a part
    // not yet resident) the part's points; the compute stream runs the
    // stages that read only the scalars (digits, sort, bucket bounds) once
    // the scalars are in, and waits for the points before the rest, so the
    // sort overlaps the larger point copy.
    u32 prev_part = ~0u;
    for (size_t o = 0; o < ops.size(); o++) {
      const Op& op = ops[o];
      const u32 k = op.pstage;
      const u64 off = (u64)op.part * h.part_len;
      const u64 cur = op.part + 1 == h.parts ? h.len - off : h.part_len;
      const u32 sk = next_sc;
      next_sc = (next_sc + 1) % h.ss;
      IMP_CHECK(hipStreamWaitEvent(g.cp, sc_free[sk], 0));
      IMP_CHECK(hipMemcpyAsync(buf + h.off_scl + sk * h.scl_stage,
                               h_scaler_batch[b0 + op.batch] + 8 * off, (size_t)cur * 32,
                               hipMemcpyHostToDevice, g.cp));
      IMP_CHECK(hipEventRecord(sc_ready[sk], g.cp));
      const bool new_part = op.part != prev_part;
      if (new_part && resident[k] != (long long)op.part) {
        // the stage's previous part is done once its last MSM has run
        IMP_CHECK(hipStreamWaitEvent(g.cp, pt_free[k], 0));
        u32* dst = (u32*)(buf + h.off_pts + k * h.pts_stage);
        for (u32 t = 0; t < h.s.n_pre; t++)
          IMP_CHECK(hipMemcpyAsync(dst + (size_t)16 * t * h.part_len, h_points[t] + 16 * off,
                                   (size_t)cur * 64, hipMemcpyHostToDevice, g.cp));
        if (h.s.n_pre > 1) {
          // table format, in place; rows past a short last part unused
          for (u32 t = 0; t < h.s.n_pre; t++)
            IMP_CHECK(to_table(dst + (size_t)16 * t * h.part_len,
                               dst + (size_t)16 * t * h.part_len, cur, g.cp));
        }
        IMP_CHECK(hipEventRecord(pt_ready[k], g.cp));
        resident[k] = op.part;
        next_pt = (k + 1) % h.sp;
      }
      IMP_CHECK(hipStreamWaitEvent(g.st, sc_ready[sk], 0));
      char* d = desc + o * h.op_desc;
      const u32 sh = op.part + 1 == h.parts ? 1 : 0;
      const std::vector<StageLaunch>& sl = stl[sh][mode[o]];
      const u32* pf = (const u32*)(d + align256(sizeof(MsmJob)));
      // the part's points are in place for every later operation on it (the
      // compute stream waited for them at the part's first operation)
      const u32 split = new_part ? pat[sh][mode[o]] : 0;
      IMP_CHECK(launch_stages((const MsmJob*)d, 1, pf, sl.data(), split, 0, g.st));
      if (new_part) {
        IMP_CHECK(hipStreamWaitEvent(g.st, pt_ready[k], 0));
        prev_part = op.part;
      }
      IMP_CHECK(launch_stages((const MsmJob*)d, 1, pf + (size_t)split * 2, sl.data() + split,
                              (u32)sl.size() - split, 0, g.st));
      IMP_CHECK(hipEventRecord(sc_free[sk], g.st));
      const bool last_of_part = o + 1 == ops.size() || ops[o + 1].part != op.part;
      if (last_of_part) IMP_CHECK(hipEventRecord(pt_free[k], g.st));
    }
    u32* fin = (u32*)(buf + h.off_fin);
    for (u32 j = 0; j < nb; j++)
      IMP_CHECK(hipMemcpyAsync(h_result[b0 + j], fin + 16 * j, 64, hipMemcpyDeviceToHost, g.st));
    IMP_CHECK(hipStreamSynchronize(g.st));
  }
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
                              unsigned long long len, unsigned int batch_per_run,
                              unsigned int parts, unsigned int stage_scalers,
                              unsigned int stage_points, unsigned int num_cards,
                              const unsigned int* cards, unsigned int const* const* h_points,
                              unsigned int batches, unsigned int const* const* h_scaler_batch,
                              unsigned int* const* h_result) {
  return msm_host(window, target_window, buffers, buffer_sizes, len, batch_per_run, parts,
                  stage_scalers, stage_points, num_cards, cards, h_points, batches,
                  h_scaler_batch, h_result);
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
  return msm_ws_total(msm_shape(len, window, 0));
}

extern "C" hipError_t impeller_msm(const uint32_t* bases_d, const uint32_t* scalars_d,
                                   unsigned long long len, unsigned int window,
                                   uint32_t* result_d, void* workspace, size_t workspace_bytes,
                                   hipStream_t stream) {
  IMP_REQUIRE(valid_window(window), "window");
  IMP_REQUIRE(result_d && workspace && (len == 0 || (bases_d && scalars_d)), "null pointer");
  MsmShape s = msm_shape(len, window, 0);
  IMP_REQUIRE((u64)s.n_pre * len < (1ull << 31), "point index range (n_pre * len)");
  IMP_REQUIRE(workspace_bytes >= msm_ws_total(s), "workspace too small");
  return msm_run(bases_d, len, scalars_d, s, result_d, (char*)workspace, stream);
}

extern "C" size_t impeller_msm_cfg_workspace_size(unsigned long long len, unsigned int window,
                                                  unsigned int target_windows,
                                                  unsigned int task) {
  if (!valid_window(window) || task == 0) return 0;
  return msm_ws_total(msm_shape(len, window, target_windows, task));
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
  MsmShape s = msm_shape(len, window, target_windows);
  if (s.n_pre <= 1) return hipSuccess;  // the plain bases: converted by each MSM call
  IMP_CHECK(precompute_device(tables_d, len, s, stream));
  return to_table(tables_d, tables_d, (u64)s.n_pre * len, stream);
}

extern "C" hipError_t impeller_msm_cfg(const uint32_t* tables_d, const uint32_t* scalars_d,
                                       unsigned long long len, unsigned long long table_stride,
                                       unsigned int window,
                                       unsigned int target_windows, unsigned int task,
                                       uint32_t* result_d, void* workspace,
                                       size_t workspace_bytes, hipStream_t stream) {
  IMP_REQUIRE(valid_window(window) && task > 0 && task <= 4096, "window / task");
  IMP_REQUIRE(result_d && workspace && (len == 0 || (tables_d && scalars_d)), "null pointer");
  MsmShape s = msm_shape(len, window, target_windows, task);
  IMP_REQUIRE(workspace_bytes >= msm_ws_total(s), "workspace too small");
  u64 stride = table_stride ? table_stride : len;
  IMP_REQUIRE(stride >= len && (u64)s.n_pre * stride < (1ull << 31), "table stride");
  return msm_run(tables_d, stride, scalars_d, s, result_d, (char*)workspace, stream);
}

// ------------------------------------------------------------ batched MSM
static MsmShape item_shape(const ImpMsmItem& it) {
  return msm_shape(it.len, it.window, it.target_windows, it.task ? it.task : 64);
}

extern "C" size_t impeller_msm_batch_bytes(const ImpMsmItem* items, unsigned int count) {
  if (!items || count == 0) return 0;
  size_t total = desc_bytes(count);
  for (u32 i = 0; i < count; i++) {
    if (!valid_window(items[i].window)) return 0;
    total += align256(msm_layout(item_shape(items[i])).total);
  }
  return total;
}

extern "C" hipError_t impeller_msm_batch_prepare(const ImpMsmItem* items, unsigned int count,
                                                 void* dev, size_t bytes, ImpMsmBatch* out,
                                                 hipStream_t stream) {
  IMP_REQUIRE(items && out && dev && count > 0, "null pointer / empty batch");
  size_t need = impeller_msm_batch_bytes(items, count);
  IMP_REQUIRE(need > 0 && bytes >= need, "batch buffer too small / bad item");
  std::vector<MsmShape> shapes;
  std::vector<MsmJob> jobs;
  char* base = (char*)dev;
  size_t off = desc_bytes(count);
  for (u32 i = 0; i < count; i++) {
    const ImpMsmItem& it = items[i];
    IMP_REQUIRE(it.result && (it.len == 0 || (it.tables && it.scalars)), "batch item");
    MsmShape s = item_shape(it);
    u64 stride = it.table_stride ? it.table_stride : it.len;
    IMP_REQUIRE(stride >= it.len, "table stride");
    IMP_REQUIRE((u64)s.n_pre * stride < (1ull << 31), "point index range");
    shapes.push_back(s);
    jobs.push_back(make_msm_job(s, stride, it.tables, it.scalars, it.result, base + off));
    off += align256(msm_layout(s).total);
  }
  std::vector<StageLaunch> st;
  std::vector<u32> prefix;
  plan_stages(shapes, st, prefix);
  IMP_REQUIRE(st.size() <= kMaxStages && st.size() <= sizeof(out->kind) / sizeof(out->kind[0]),
              "too many stages");
  out->dev = dev;
  out->count = count;
  out->nstages = (unsigned int)st.size();
  for (size_t k = 0; k < st.size(); k++) {
    out->kind[k] = st[k].kind;
    out->arg[k] = st[k].arg;
    out->total[k] = st[k].total;
  }
  IMP_CHECK(hipMemcpyAsync(base, jobs.data(), sizeof(MsmJob) * count, hipMemcpyHostToDevice,
                           stream));
  IMP_CHECK(hipMemcpyAsync(base + align256(sizeof(MsmJob) * count), prefix.data(),
                           prefix.size() * 4, hipMemcpyHostToDevice, stream));
  return hipStreamSynchronize(stream);
}

extern "C" hipError_t impeller_msm_batch_run(const ImpMsmBatch* b, unsigned int max_blocks,
                                             hipStream_t stream) {
  IMP_REQUIRE(b && b->dev && b->count > 0, "batch");
  const char* base = (const char*)b->dev;
  const MsmJob* jobs = (const MsmJob*)base;
  const u32* prefix = (const u32*)(base + align256(sizeof(MsmJob) * b->count));
  std::vector<StageLaunch> st(b->nstages);
  for (u32 k = 0; k < b->nstages; k++) st[k] = StageLaunch{b->kind[k], b->arg[k], b->total[k]};
  return launch_stages(jobs, b->count, prefix, st.data(), b->nstages, max_blocks, stream);
}
