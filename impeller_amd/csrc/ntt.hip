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
// Passes (in place, no workspace, unless the bit reversal is fused):
//   1. bit-reversal permutation: tile pairs (b, rev b) of 16 x 16 (or 32 x 32)
//      elements swapped through LDS, optionally multiplying x[q] by shift^q
//      (coset) and running the first 4 (or 5) DIT levels; or, fused into the
//      first LDS pass, reads of the source at bit-reversed positions (that
//      pass then writes a second buffer, which the next pass reads back);
//   2. ceil(L / 8) DIT passes; each work-group stages a tile of G = 2^D
//      positions x C columns in LDS and runs D consecutive levels, columns
//      chosen so that every global access is a run of >= 4 consecutive
//      32-byte elements (one 128-byte line);
//   3. the last pass can multiply each output by post(i) (n^-1 and/or
//      shift^-i for inverse and inverse-coset transforms).
#include <mutex>
#include <vector>

#include "common.h"
#include "host_field.h"
#include "timing.h"

namespace imp {

// --------------------------------------------------------- on-the-fly factors
// factor(i) = T0[i mod 2^h] * T1[(i >> h) mod 2^h] * T2[i >> 2h] with h-bit
// digits, h = ceil(L / 2) (at most kFactorDigit): one table for L <= 13,
// one multiply for L <= 26, two above.  The tables hold exact powers, so the
// product equals the direct power bit for bit.
static const u32 kFactorDigit = 13;
static const u32 kFactorTab = 1u << kFactorDigit;  // entries per digit table

struct Factor {
  const u32* tab;  // 3 * kFactorTab elements, or nullptr (factor 1)
  u32 h;           // digit bits
  u32 ndig;        // digits used (1..3)
};

__host__ __device__ inline Factor factor_shape(const u32* tab, u32 L) {
  Factor f;
  f.tab = tab;
  if (L <= kFactorDigit) {
    f.h = L ? L : 1;
    f.ndig = 1;
  } else if (L <= 2 * kFactorDigit) {
    f.h = (L + 1) / 2;
    f.ndig = 2;
  } else {
    f.h = kFactorDigit;
    f.ndig = 3;
  }
  return f;
}

__device__ __forceinline__ Fr factor_at(const Factor& f, u64 i) {
  const u64 mask = (1ull << f.h) - 1;
  Fr r = load_fe<Fr>(f.tab + 8 * (i & mask));
  if (f.ndig > 1)
    r = mul<FrParams>(r, load_fe<Fr>(f.tab + 8 * (kFactorTab + ((i >> f.h) & mask))));
  if (f.ndig > 2)
    r = mul<FrParams>(r, load_fe<Fr>(f.tab + 8 * (2 * kFactorTab + (i >> (2 * f.h)))));
  return r;
}

// tab[d * kFactorTab + j] = scale^(d==0) * base^(j * 2^(h d)), j < 2^h
__global__ void k_factor_tables(u32* tab, const u32* base_p, const u32* scale_p, u32 h,
                                u32 ndig) {
  u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 per = 1u << h;
  if (t >= per * ndig) return;
  u32 d = t / per, j = t % per;
  Fr base = base_p ? load_fe<Fr>(base_p) : Fr::one();
  for (u32 k = 0; k < h * d; k++) base = sqr<FrParams>(base);
  Fr r = pow_u32<FrParams>(base, j);
  if (d == 0 && scale_p) r = mul<FrParams>(r, load_fe<Fr>(scale_p));
  store_fe<Fr>(tab + 8 * ((u64)d * kFactorTab + j), r);
}

static hipError_t launch_factor_tables(const Factor& f, const u32* base_d, const u32* scale_d,
                                       hipStream_t stream) {
  const u32 n = (1u << f.h) * f.ndig;
  hipLaunchKernelGGL(k_factor_tables, dim3((n + 255) / 256), dim3(256), 0, stream,
                     const_cast<u32*>(f.tab), base_d, scale_d, f.h, f.ndig);
  return hipGetLastError();
}

// --------------------------------------------------------------- twiddles
// Direct mode: tab[e] = omega^e for e < n/2 (the caller's table).
// Recompute mode (recompute_ntt, for callers without room for that table):
// pq[i] = omega^(i 2^(L - pq_deg)) for i < 2^(pq_deg - 1) and omegas[k] =
// omega^(2^k); every LDS-pass work-group derives the twiddles of its own tile
// from them into LDS before its butterflies (rc_prologue), so the transform
// needs no memory beyond pq and omegas.  Field products are exact, so these
// twiddles equal the table's entries bit for bit.
struct Twiddle {
  const u32* tab;
  const u32* pq;      // recompute mode; nullptr: everything from omegas
  const u32* omegas;  // recompute mode
  u32 pq_deg;
};

// ------------------------------------------------------ butterfly helpers
// radix-2: (a, b) <- (a + w b, a - w b); e == 0 means w = 1
__device__ __forceinline__ Fr tw_load(const Twiddle& tw, u64 e) {
  return load_fe<Fr>(tw.tab + 8 * e);
}

// omega^e as the product of omegas[k] over the set bits k of e (e < 2^32)
__device__ __forceinline__ Fr pow_omegas(const u32* omegas, u64 e) {
  Fr acc = Fr::one();
  bool first = true;
  for (u32 k = 0; e; k++, e >>= 1) {
    if (e & 1) {
      const Fr w = load_fe<Fr>(omegas + 8 * k);
      acc = first ? w : mul<FrParams>(acc, w);
      first = false;
    }
  }
  return acc;
}

// Lazy butterflies of the LDS passes (Harvey, "Faster arithmetic for
// number-theoretic transforms", 2014): values live in [0, 4m) between levels
// and between passes (4m < 2^256 for Fr; the last pass stores canonical values).  (a, b) -> (a' + t, a' + 2m - t) with a' = a mod 2m
// folded into [0, 2m) and t = b*w by the Montgomery product without its final
// subtraction (b < 4m, w < m give t < 2m), or b folded into [0, 2m) when w = 1.
__device__ __forceinline__ Fr fold2m(const Fr& a) {  // [0, 4m) -> [0, 2m)
  Fr d;
  u32 br = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const u32 m2 = (FrParams::m[i] << 1) | (i ? FrParams::m[i - 1] >> 31 : 0u);
    d.v[i] = sbb(a.v[i], m2, br, &br);
  }
  Fr r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.v[i] = br ? a.v[i] : d.v[i];
  return r;
}

__device__ __forceinline__ void bfly2_lazy(Fr& a, Fr& b, bool m, const Fr& w) {
  const Fr x = fold2m(a);
  const Fr t = m ? mul_lazy<FrParams>(b, w) : fold2m(b);
  Fr u;  // 2m - t, t < 2m
  u32 br = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const u32 m2 = (FrParams::m[i] << 1) | (i ? FrParams::m[i - 1] >> 31 : 0u);
    u.v[i] = sbb(m2, t.v[i], br, &br);
  }
  u32 c0 = 0, c1 = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    a.v[i] = adc(x.v[i], t.v[i], c0, &c0);
    b.v[i] = adc(x.v[i], u.v[i], c1, &c1);
  }
}

// [0, 4m) -> [0, m)
__device__ __forceinline__ Fr canon4m(const Fr& a) { return reduce_once<FrParams>(fold2m(a)); }

__device__ __forceinline__ void bfly2_v(Fr& a, Fr& b, u64 e, const Twiddle& tw) {
  if (e != 0) b = mul<FrParams>(b, tw_load(tw, e));
  Fr s = add<FrParams>(a, b);
  b = sub<FrParams>(a, b);
  a = s;
}

// Two consecutive DIT levels on positions g0 < g1 = g0 + 2^j < g2 = g0 +
// 2^(j+1) < g3: level j pairs (g0,g1),(g2,g3) with exponent e1; level j+1
// pairs (g0,g2) with e2 and (g1,g3) with e3.  The same butterflies, in the
// same order per element, as two radix-2 levels.
__device__ __forceinline__ void bfly4_v(Fr& a0, Fr& a1, Fr& a2, Fr& a3, u64 e1, u64 e2, u64 e3,
                                        const Twiddle& tw) {
  if (e1 != 0) {
    Fr w1 = tw_load(tw, e1);
    a1 = mul<FrParams>(a1, w1);
    a3 = mul<FrParams>(a3, w1);
  }
  Fr b0 = add<FrParams>(a0, a1), b1 = sub<FrParams>(a0, a1);
  Fr b2 = add<FrParams>(a2, a3), b3 = sub<FrParams>(a2, a3);
  if (e2 != 0) b2 = mul<FrParams>(b2, tw_load(tw, e2));
  b3 = mul<FrParams>(b3, tw_load(tw, e3));
  a0 = add<FrParams>(b0, b2);
  a2 = sub<FrParams>(b0, b2);
  a1 = add<FrParams>(b1, b3);
  a3 = sub<FrParams>(b1, b3);
}

// Planar LDS tiles: slot s of a buffer of ns slots keeps its low 16 bytes at
// u32 offset 4 s and its high 16 bytes at 4 (ns + s), so the 16-byte accesses
// of 16 consecutive lanes hit distinct banks whenever their slots differ
// mod 16 (a 32-byte element stored whole would span two bank groups).
struct Planar {
  u32* p;
  u32 ns;
};

__device__ __forceinline__ Fr pl_load(const Planar& t, u32 s) {
  const uint4* q = reinterpret_cast<const uint4*>(t.p);
  uint4 a = q[s], b = q[t.ns + s];
  Fr r;
  r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
  r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
  return r;
}

__device__ __forceinline__ void pl_store(const Planar& t, u32 s, const Fr& v) {
  uint4* q = reinterpret_cast<uint4*>(t.p);
  q[s] = make_uint4(v.v[0], v.v[1], v.v[2], v.v[3]);
  q[t.ns + s] = make_uint4(v.v[4], v.v[5], v.v[6], v.v[7]);
}

// ------------------------------------------------------------ job descriptor
// One transform = a bit-reversal stage (optionally with the first 5 DIT
// levels) and nst LDS passes.  src is read by the bit-reversal stage (the
// same view as x for an in-place transform); source positions >= src_len
// read as zero (zero-extension of a 2^k polynomial to 2^(k+e)).
struct NttStage {
  u32 s, D, logC;
  u32 blocks;
};

struct NttJob {
  View x;
  View src;
  u64 src_len;
  u32 L;
  u32 fuse;  // DIT levels fused into the bit reversal: 0, or tb (L >= 2 tb)
  u32 tb;    // bit-reversal tile side 2^tb (4: 16 x 16, two tile pairs per
             // 256-thread work-group; 5: 32 x 32, one pair)
  u32 rc;    // recompute-mode twiddles (then fuse == 0 and no fused reversal)
  Twiddle tw;
  Factor pro, epi;  // tab == nullptr: none
  u32 nst;
  u32 br_blocks;
  u32 fbr;   // bit reversal fused into stage 0 (out of place through tmp)
  u32 zlev;  // fbr only: source positions >= 2^(L - zlev) read as zero, so
             // stage 0's first zlev levels are copies (zlev <= its depth)
  u32* tmp;  // 2^L elements: stage 0 writes, stage 1 reads (fbr only)
  NttStage st[8];
};

__device__ __forceinline__ Fr src_load(const NttJob& J, u64 q) {
  if (q >= J.src_len) return Fr::zero();
  Fr v = load_fe<Fr>(J.src.base + 8 * view_phys(J.src, q));
  if (J.pro.tab) v = mul<FrParams>(v, factor_at(J.pro, q));
  return v;
}

// Bit reversal on tile pairs with tiles of NT x NT = 2^TB x 2^TB elements:
// q = (a | b | c) (TB | L-2TB | TB bits) moves to p = rev(q) = (rev c | rev b |
// rev a), i.e. from tile b to tile rev(b); the pair is owned by one group of P
// threads, so the permutation runs in place.  In LDS a destination tile is
// row-major by p's high TB bits (row) and low TB bits (column), the column
// XOR-swizzled by the row (planar storage, conflict-free for the transposing
// writes, the row-wise butterflies and the row-wise reads).  Fused DIT levels
// 0..TB-1 pair up columns: level j has exponent (col mod 2^j) 2^(L-j-1).
// Every thread of the work-group reaches every barrier (valid/active only
// gate the work), since two pairs can share one work-group.
template <u32 TB>
__device__ __forceinline__ void br_tile_pair(const NttJob& J, u64 b, bool valid, u32 tid, u32 P,
                                             const Planar& T) {
  const u32 L = J.L, NT = 1u << TB, TS = NT * NT;
  const u64 rb = valid ? bit_reverse(b, L - 2 * TB) : 0;
  const bool active = valid && b <= rb;
  const bool self = b == rb;
  __syncthreads();
  if (active) {
    for (u32 idx = tid; idx < TS; idx += P) {
      const u32 a = idx >> TB, c = idx & (NT - 1);
      const u32 row = brev_bits(c, TB), col = brev_bits(a, TB);
      const u32 slot = row * NT + (col ^ row);
      pl_store(T, (self ? 0 : TS) + slot, src_load(J, ((u64)a << (L - TB)) | (b << TB) | c));
      if (!self) pl_store(T, slot, src_load(J, ((u64)a << (L - TB)) | (rb << TB) | c));
    }
  }
  __syncthreads();
  if (J.fuse == TB) {
    const u32 nt = self ? 1 : 2;
    for (u32 j = 0; j + 1 < TB; j += 2) {
      if (active) {
        for (u32 qi = tid; qi < nt * (TS / 4); qi += P) {
          const u32 tl = qi / (TS / 4), q = qi % (TS / 4);
          const u32 row = q / (NT / 4), k = q % (NT / 4);
          const u32 glo = k & ((1u << j) - 1);
          const u32 c0 = ((k >> j) << (j + 2)) | glo;
          const u32 rs = tl * TS + row * NT;
          const u32 s0 = rs + (c0 ^ row), s1 = rs + ((c0 | (1u << j)) ^ row);
          const u32 s2 = rs + ((c0 | (2u << j)) ^ row), s3 = rs + ((c0 | (3u << j)) ^ row);
          Fr a0 = pl_load(T, s0), a1 = pl_load(T, s1), a2 = pl_load(T, s2), a3 = pl_load(T, s3);
          bfly4_v(a0, a1, a2, a3, (u64)glo << (L - j - 1), (u64)glo << (L - j - 2),
                  ((u64)glo + (1u << j)) << (L - j - 2), J.tw);
          pl_store(T, s0, a0);
          pl_store(T, s1, a1);
          pl_store(T, s2, a2);
          pl_store(T, s3, a3);
        }
      }
      __syncthreads();
    }
    if (TB & 1) {
      const u32 j = TB - 1;
      if (active) {
        for (u32 bi = tid; bi < nt * (TS / 2); bi += P) {
          const u32 tl = bi / (TS / 2), q = bi % (TS / 2);
          const u32 row = q / (NT / 2), k = q % (NT / 2);
          const u32 rs = tl * TS + row * NT;
          const u32 s0 = rs + (k ^ row), s1 = rs + ((k | (1u << j)) ^ row);
          Fr a0 = pl_load(T, s0), a1 = pl_load(T, s1);
          bfly2_v(a0, a1, (u64)k << (L - TB), J.tw);
          pl_store(T, s0, a0);
          pl_store(T, s1, a1);
        }
      }
      __syncthreads();
    }
  }
  if (active) {
    for (u32 idx = tid; idx < TS; idx += P) {
      const u32 row = idx >> TB, col = idx & (NT - 1);
      const u32 slot = row * NT + (col ^ row);
      const u64 p0 = ((u64)row << (L - TB)) | (b << TB) | col;
      store_fe<Fr>(J.x.base + 8 * view_phys(J.x, p0), pl_load(T, slot));
      if (!self) {
        const u64 p1 = ((u64)row << (L - TB)) | (rb << TB) | col;
        store_fe<Fr>(J.x.base + 8 * view_phys(J.x, p1), pl_load(T, TS + slot));
      }
    }
  }
}

// whole array (L < 2 tb) in one work-group
__device__ __forceinline__ void br_small_body(const NttJob& J, u32* lds) {
  const u32 n = 1u << J.L;
  __syncthreads();
  for (u32 i = threadIdx.x; i < n; i += blockDim.x)
    store_fe<Fr>(lds + 8 * brev_bits(i, J.L), src_load(J, i));
  __syncthreads();
  for (u32 i = threadIdx.x; i < n; i += blockDim.x)
    store_fe<Fr>(J.x.base + 8 * view_phys(J.x, i), load_fe<Fr>(lds + 8 * i));
}

// block lb of the bit-reversal stage (256 threads)
__device__ __forceinline__ void br_body(const NttJob& J, u64 lb, u32* lds) {
  if (J.L < 2 * J.tb) {
    br_small_body(J, lds);
  } else if (J.tb == 5) {
    br_tile_pair<5>(J, lb, true, threadIdx.x, 256, Planar{lds, 2 * 1024});
  } else {
    const u32 half = threadIdx.x >> 7;
    const u64 b = 2 * lb + half;
    br_tile_pair<4>(J, b, b < (1ull << (J.L - 8)), threadIdx.x & 127, 128,
                    Planar{lds + half * (2 * 256 * 8), 2 * 256});
  }
}

// 16-byte slot swizzle of the DIT tiles (planar storage): the low four slot
// bits are XORed with the next eight, so the power-of-two slot strides of the
// butterfly levels and of the load / store phases spread over the 16 slot
// groups of a bank row; a bijection on every power-of-two range >= 16.
__device__ __forceinline__ u32 dslot(u32 s) { return s ^ (((s >> 4) ^ (s >> 8)) & 15u); }

// LDS pass over levels s .. s+D-1 on positions p = hi*2^(s+D) + g*2^s + lo.
// contiguous (s == 0): work-group = C consecutive groups of G positions.
// strided (2^s >= C): work-group = fixed hi, C consecutive lo.  The last pass
// multiplies each output by the epilogue factor (n^-1, coset un-shift).
// RC: recompute-mode twiddles (J.rc), derived per work-group into LDS past
// the tile.  Each butterfly issues its twiddle loads before its LDS reads, so
// one wait covers both.
//
// With the bit reversal fused (J.fbr), stage 0 takes for work-group m the
// source positions q = (g' << (L-D)) | (m << logC) | c (runs of C elements),
// i.e. the transform positions p = rev(q) = (rev c << (L-logC)) |
// (rev m << D) | rev g': tile row g = rev g' of column c, the same tile as a
// contiguous pass over those columns.  It writes p to J.tmp (runs of G
// elements); stage 1 reads J.tmp and writes x.
// Recompute-mode twiddles of one work-group (levels s .. s+D-1, columns lo =
// lo0 + c; lo = 0 in a contiguous pass).  The twiddle of level s + j, group
// offset glo < 2^j, column c is omega^e with e = (glo 2^s + lo) 2^(L-s-j-1)
// = glo 2^(L-j-1) + lo 2^(L-s-j-1), the product of
//   P[glo << (D-1-j)],  P[k] = omega^(k 2^(L-D)) for k < 2^(D-1)
//                       (= pq[k << (pq_deg - D)] when D <= pq_deg <= L), and
//   Q[c D + j] = omega^(lo 2^(L-s-j-1)) (1 for lo = 0),
// both left in LDS at P = tile end, Q = P + 8 2^(D-1).
__device__ __forceinline__ void rc_prologue(const NttJob& J, u32 L, u32 s, u32 D, u32 C,
                                            bool contig, u64 lo0, u32* P) {
  const u32 np = 1u << (D - 1);
  const bool use_pq = J.tw.pq != nullptr && J.tw.pq_deg >= D && J.tw.pq_deg <= L;
  for (u32 k = threadIdx.x; k < np; k += blockDim.x) {
    const Fr v = use_pq ? load_fe<Fr>(J.tw.pq + 8 * ((u64)k << (J.tw.pq_deg - D)))
                        : pow_omegas(J.tw.omegas, (u64)k << (L - D));
    store_fe<Fr>(P + 8 * k, v);
  }
  u32* Q = P + 8 * np;
  for (u32 q = threadIdx.x; q < C * D; q += blockDim.x) {
    const u32 c = q / D, j = q % D;
    const u64 lo = contig ? 0 : lo0 + c;
    store_fe<Fr>(Q + 8 * q, pow_omegas(J.tw.omegas, lo << (L - s - j - 1)));
  }
}

__device__ __forceinline__ Fr rc_twiddle(const u32* P, const u32* Q, u32 k, u32 q, bool qm) {
  const Fr w = load_fe<Fr>(P + 8 * k);
  return qm ? mul<FrParams>(w, load_fe<Fr>(Q + 8 * q)) : w;
}

template <bool RC>
__device__ __forceinline__ void dit_body(const NttJob& J, u32 stage, u64 wg, u32* t) {
  const NttStage& S = J.st[stage];
  const bool last = stage + 1 == J.nst;
  const u32 L = J.L, s = S.s, D = S.D, logC = S.logC;
  const u32 C = 1u << logC, G = 1u << D, GC = G * C;
  const bool contig = (s == 0);
  const bool rd_rev = J.fbr && stage == 0, rd_tmp = J.fbr && stage == 1;
  u64 hi = 0, lo0 = 0;
  if (!contig) {
    // strided: the twiddles depend on lo only.  Work-groups are dealt to the
    // 8 XCDs round-robin, so let each XCD take whole lo groups (wg mod 8
    // picks the group within a set of 8) and run all hi of a group back to
    // back: the group's twiddles are then read into that XCD's L2 once.
    const u32 lg_lo = s - logC, lg_hi = L - s - D;
    if (lg_lo >= 3) {
      const u64 k = wg >> 3;
      hi = k & ((1ull << lg_hi) - 1);
      lo0 = (((k >> lg_hi) << 3) | (wg & 7)) << logC;
    } else {
      hi = wg >> lg_lo;
      lo0 = (wg & ((1ull << lg_lo) - 1)) << logC;
    }
  }
  const Planar T{t, GC};
  u32* const P = t + 8 * GC;  // recompute-mode twiddles past the tile
  const u32* const Q = P + 8 * (1u << (D - 1));
  __syncthreads();
  if (RC) rc_prologue(J, L, s, D, C, contig, lo0, P);
  // zero-extended source (zlev > 0): a row g = rev gq is nonzero only when
  // its low zlev bits are clear (gq < 2^(D - zlev)), and DIT levels
  // 0 .. zlev-1 with a zero partner (a, 0) -> (a, a) only copy it to the rows
  // that differ in those bits: store those copies and start at level zlev
  const u32 z = rd_rev ? J.zlev : 0;
  if (rd_rev) {
    for (u32 idx = threadIdx.x; idx < (GC >> z); idx += blockDim.x) {
      const u32 c = idx & (C - 1), gq = idx >> logC;
      const u64 q = ((u64)gq << (L - D)) | (wg << logC) | c;
      const Fr v = src_load(J, q);
      const u32 g = brev_bits(gq, D);
      for (u32 r = 0; r < (1u << z); r++) pl_store(T, dslot((g | r) * C + c), v);
    }
  } else {
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
      const u32* src = rd_tmp ? J.tmp + 8 * p : J.x.base + 8 * view_phys(J.x, p);
      pl_store(T, dslot(g * C + c), load_fe<Fr>(src));
    }
  }
  __syncthreads();
  u32 j = z;
  for (; j + 1 < D; j += 2) {
    const u32 l = s + j;
    const u32 sh1 = L - l - 2;
    for (u32 qi = threadIdx.x; qi < (GC >> 2); qi += blockDim.x) {
      u32 c = qi & (C - 1), k = qi >> logC;
      u32 glo = k & ((1u << j) - 1);
      u32 g0 = ((k >> j) << (j + 2)) | glo;
      u64 lo = contig ? 0 : lo0 + c;
      u64 base = (((u64)glo) << s) + lo;
      u64 e1 = base << (sh1 + 1), e2 = base << sh1, e3 = (base + (1ull << l)) << sh1;
      const u32 s0 = dslot(g0 * C + c), s1 = dslot((g0 | (1u << j)) * C + c);
      const u32 s2 = dslot((g0 | (2u << j)) * C + c), s3 = dslot((g0 | (3u << j)) * C + c);
      Fr w1, w2, w3;
      if (RC) {
        const bool qm = !contig && lo != 0;
        w1 = rc_twiddle(P, Q, glo << (D - 1 - j), c * D + j, qm);
        w2 = rc_twiddle(P, Q, glo << (D - 2 - j), c * D + j + 1, qm);
        w3 = rc_twiddle(P, Q, (glo + (1u << j)) << (D - 2 - j), c * D + j + 1, qm);
      } else {
        w1 = tw_load(J.tw, e1);
        w2 = tw_load(J.tw, e2);
        w3 = tw_load(J.tw, e3);
      }
      Fr a0 = pl_load(T, s0), a1 = pl_load(T, s1), a2 = pl_load(T, s2), a3 = pl_load(T, s3);
      // two levels: (a0, a1), (a2, a3) with w1, then (a0, a2) with w2 and
      // (a1, a3) with w3 -- best_fft's butterflies, in its order per element
      bfly2_lazy(a0, a1, e1 != 0, w1);
      bfly2_lazy(a2, a3, e1 != 0, w1);
      bfly2_lazy(a0, a2, e2 != 0, w2);
      bfly2_lazy(a1, a3, true, w3);
      pl_store(T, s0, a0);
      pl_store(T, s1, a1);
      pl_store(T, s2, a2);
      pl_store(T, s3, a3);
    }
    __syncthreads();
  }
  if (j < D) {
    const u32 l = s + j;
    const u32 sh = L - l - 1;
    for (u32 bi = threadIdx.x; bi < (GC >> 1); bi += blockDim.x) {
      u32 c = bi & (C - 1), k = bi >> logC;
      u32 glo = k & ((1u << j) - 1);
      u32 g0 = ((k >> j) << (j + 1)) | glo, g1 = g0 | (1u << j);
      u64 lo = contig ? 0 : lo0 + c;
      u64 e = ((((u64)glo) << s) + lo) << sh;
      const u32 s0 = dslot(g0 * C + c), s1 = dslot(g1 * C + c);
      const Fr w = RC ? rc_twiddle(P, Q, glo << (D - 1 - j), c * D + j, !contig && lo != 0)
                      : tw_load(J.tw, e);
      Fr a0 = pl_load(T, s0), a1 = pl_load(T, s1);
      bfly2_lazy(a0, a1, e != 0, w);
      pl_store(T, s0, a0);
      pl_store(T, s1, a1);
    }
    __syncthreads();
  }
  const bool epi = last && J.epi.tab != nullptr;
  // fused reversal: column c of work-group m holds p = (rev c << (L-logC)) |
  // (rev m << D) | g
  const u64 rev_m = rd_rev ? (u64)brev_bits((u32)wg, L - D - logC) << D : 0;
  for (u32 idx = threadIdx.x; idx < GC; idx += blockDim.x) {
    u32 g, c;
    u64 p;
    if (rd_rev) {
      c = idx >> D;
      g = idx & (G - 1);
      p = ((u64)brev_bits(c, logC) << (L - logC)) | rev_m | g;
    } else if (contig) {
      c = idx >> D;
      g = idx & (G - 1);
      p = (((wg << logC) + c) << D) | g;
    } else {
      g = idx >> logC;
      c = idx & (C - 1);
      p = (hi << (s + D)) | ((u64)g << s) | (lo0 + c);
    }
    Fr v = pl_load(T, dslot(g * C + c));
    // between passes the lazy value goes to memory as it is (the next
    // pass's butterflies take [0, 4m)); the last pass stores canonical values
    if (epi) v = mul<FrParams>(v, factor_at(J.epi, p));
    else if (last) v = canon4m(v);
    store_fe<Fr>(rd_rev ? J.tmp + 8 * p : J.x.base + 8 * view_phys(J.x, p), v);
  }
}

// --------------------------------------------------------------- kernels
__global__ __launch_bounds__(256, 4) void k_ntt_br(NttJob J) {
  IMP_DYNAMIC_LDS(lds);
  br_body(J, blockIdx.x, lds);
}

template <bool RC>
__global__ __launch_bounds__(256, 4) void k_ntt_dit(NttJob J, u32 stage) {
  IMP_DYNAMIC_LDS(lds);
  dit_body<RC>(J, stage, blockIdx.x, lds);
}

__device__ __forceinline__ u32 find_job(const u32* prefix, u32 njobs, u32 t) {
  u32 lo = 0, hi = njobs - 1;
  while (lo < hi) {
    u32 mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Batched stages: work-groups stride over the concatenated block space of all
// jobs (prefix[j] = first block of job j in this stage); jobs without work in
// the stage contribute zero blocks.
__global__ __launch_bounds__(256, 4) void k_ntt_br_batch(const NttJob* jobs, u32 njobs,
                                                      const u32* prefix) {
  IMP_DYNAMIC_LDS(lds);
  const u32 total = prefix[njobs];
  for (u32 t = blockIdx.x; t < total; t += gridDim.x) {
    u32 j = find_job(prefix, njobs, t);
    br_body(jobs[j], t - prefix[j], lds);
  }
}

__global__ __launch_bounds__(256, 4) void k_ntt_dit_batch(const NttJob* jobs, u32 njobs,
                                                       const u32* prefix, u32 stage) {
  IMP_DYNAMIC_LDS(lds);
  const u32 total = prefix[njobs];
  for (u32 t = blockIdx.x; t < total; t += gridDim.x) {
    u32 j = find_job(prefix, njobs, t);
    // batch jobs always carry a twiddle table (never recompute mode)
    dit_body<false>(jobs[j], stage, t - prefix[j], lds);
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

// n^-1 for n = 2^L: Fermat inversion of from_number(2^L)
__global__ void k_ninv(u32* out, u32 L) {
  Fr n = Fr::zero();
  n.v[L / 32] = 1u << (L % 32);
  n = from_number<FrParams>(n);
  store_fe<Fr>(out, inv<FrParams>(n));
}

// ----------------------------------------------------------------- planner
// fuse: DIT levels done inside the bit-reversal kernel (0; 4 with 16 x 16 tiles,
// needs L >= 8; 5 with 32 x 32 tiles, needs L >= 10), or 1: no bit-reversal
// kernel, the first LDS pass reads the source bit-reversed (needs >= 2 passes);
// lds_log: log2 of the elements one work-group stages (10: 32 KiB, 11: 64 KiB);
// d[]: level counts of the following LDS passes.
struct PassPlan {
  u32 fuse;
  u32 lds_log;
  u32 n;
  u32 d[8];
};

// Default with a second buffer: from 2^20 up, the bit reversal fused into a
// first pass of 8 levels (runs of 4 elements) and the rest in passes of at
// most 9 levels (measured best or within a few percent of it for 2^20 ..
// 2^26); below, 4 levels in the bit-reversal kernel and passes of at most 8.
static void default_plan(u32 L, PassPlan* p) {
  p->lds_log = 10;
  if (L >= 20) {
    p->fuse = 1;
    const u32 rest = L - 8, np = (rest + 8) / 9;
    p->n = 1 + np;
    p->d[0] = 8;
    for (u32 i = 0; i < np; i++) p->d[1 + i] = rest / np + (i < rest % np ? 1 : 0);
    return;
  }
  p->fuse = L >= 8 ? 4 : 0;
  u32 rest = L - p->fuse;
  u32 np = (rest + 7) / 8;
  if (np == 0) np = 1;
  p->n = np;
  for (u32 i = 0; i < np; i++) p->d[i] = rest / np + (i < rest % np ? 1 : 0);
}

// Default in place (no second buffer: the reference-ABI ssip_ntt, and any
// call without room for one).  The fastest in-place plans the tuner measured
// on MI355X: 2^20 .. 2^23 four levels in the bit-reversal kernel and two or
// three passes; from 2^24 a plain bit reversal, a first pass of 10 levels
// (1024 consecutive elements per group) and passes of at most 8.
static void default_inplace_plan(u32 L, PassPlan* p) {
  static const u32 mid[4][4] = {{2, 8, 8, 0}, {2, 9, 8, 0}, {3, 6, 6, 6}, {3, 7, 6, 6}};
  if (L < 20) {
    default_plan(L, p);
    return;
  }
  p->lds_log = 10;
  if (L <= 23) {
    p->fuse = 4;
    p->n = mid[L - 20][0];
    for (u32 i = 0; i < p->n; i++) p->d[i] = mid[L - 20][1 + i];
    return;
  }
  p->fuse = 0;
  p->d[0] = 10;
  u32 rest = L - 10, n = 1;
  while (rest > 0 && n < 8) {
    const u32 d = rest > 8 ? 8 : rest;
    p->d[n++] = d;
    rest -= d;
  }
  p->n = n;
}

// recompute_ntt: a plain bit reversal (its kernel never needs twiddles) and
// passes of at most 8 levels, split evenly
static void recompute_plan(u32 L, PassPlan* p) {
  p->lds_log = 10;
  p->fuse = 0;
  u32 np = (L + 7) / 8;
  if (np == 0) np = 1;
  p->n = np;
  for (u32 i = 0; i < np; i++) p->d[i] = L / np + (i < L % np ? 1 : 0);
}

// plan array: {fuse, lds_log, d0, d1, ...}; fuse + sum(d) == L
static bool plan_from(const unsigned int* plan, unsigned int n, u32 L, PassPlan* p) {
  if (!plan || n == 0) {
    default_plan(L, p);
    return true;
  }
  if (n < 3 || n > 10) return false;
  p->fuse = plan[0];
  p->lds_log = plan[1];
  if (!(p->fuse == 0 || (p->fuse == 1 && n >= 4) || (p->fuse == 4 && L >= 8) ||
        (p->fuse == 5 && L >= 10)))
    return false;
  if (p->lds_log != 10 && p->lds_log != 11) return false;
  u32 sum = p->fuse == 1 ? 0 : p->fuse;
  p->n = n - 2;
  for (u32 i = 0; i < p->n; i++) {
    u32 d = plan[2 + i];
    if (d == 0 || d > p->lds_log) return false;
    p->d[i] = d;
    sum += d;
  }
  return sum == L;
}

static size_t stage_lds(const NttJob& J, u32 stage) {
  const NttStage& st = J.st[stage];
  size_t b = ((size_t)32 << st.D) << st.logC;
  // recompute mode: P (2^(D-1) twiddles) and Q (C x D column factors)
  if (J.rc) b += (size_t)32 * ((1u << (st.D - 1)) + (st.D << st.logC));
  return b;
}
static size_t br_lds(const NttJob& J) {
  if (J.L < 2 * J.tb) return (size_t)32 << J.L;
  return J.tb == 5 ? (size_t)2 * 1024 * 32 : (size_t)2 * 2 * 256 * 32;
}

static NttJob make_job(View x, View src, u64 src_len, u32 L, const Twiddle& tw, bool rc,
                       Factor pro, Factor epi, const PassPlan& plan) {
  NttJob J;
  J.x = x;
  J.src = src;
  J.src_len = src_len;
  J.L = L;
  J.tb = plan.fuse == 5 ? 5 : 4;
  J.fuse = (plan.fuse == J.tb && L >= 2 * J.tb) ? J.tb : 0;
  J.fbr = plan.fuse == 1 ? 1 : 0;
  J.zlev = 0;
  J.tmp = nullptr;
  J.rc = rc ? 1 : 0;
  J.tw = tw;
  J.pro = pro;
  J.epi = epi;
  if (L < 2 * J.tb) J.br_blocks = 1;
  else J.br_blocks = J.tb == 5 ? (u32)(1ull << (L - 10)) : (u32)(L > 8 ? 1ull << (L - 9) : 1);
  J.nst = 0;
  u32 s = J.fuse;
  for (u32 pi = 0; pi < plan.n && L > 0; pi++) {
    u32 D = plan.d[pi];
    if (D == 0) continue;
    u32 maxLogC = D >= plan.lds_log ? 0 : plan.lds_log - D;
    u32 logC = s == 0 ? (maxLogC < L - D ? maxLogC : L - D) : (maxLogC < s ? maxLogC : s);
    NttStage& st = J.st[J.nst++];
    st.s = s;
    st.D = D;
    st.logC = logC;
    st.blocks = (u32)(1ull << (L - D - logC));
    s += D;
  }
  if (J.fbr && J.nst > 0) {
    // zero extension: the source fits in 2^(L - e) positions
    u32 lg = 0;
    while (lg < L && (1ull << lg) < src_len) lg++;
    const u32 e = L - lg;
    J.zlev = e < J.st[0].D ? e : J.st[0].D;
  }
  return J;
}

// A fused-reversal plan without a second buffer runs the same passes in
// place after the bit-reversal kernel (same result).
static void drop_fused_reversal(NttJob* J) {
  J->fbr = 0;
  J->zlev = 0;
  J->tmp = nullptr;
}

static hipError_t launch_stages(const NttJob& J, hipStream_t stream) {
  if (!J.fbr) {
    TimedScope ts(K_NTT_BITREV, stream);
    hipLaunchKernelGGL(k_ntt_br, dim3(J.br_blocks), dim3(256), br_lds(J), stream, J);
    IMP_CHECK(hipGetLastError());
  }
  for (u32 st = 0; st < J.nst; st++) {
    TimedScope ts(K_NTT_DIT, stream);
    if (J.rc)
      hipLaunchKernelGGL(k_ntt_dit<true>, dim3(J.st[st].blocks), dim3(256), stage_lds(J, st),
                         stream, J, st);
    else
      hipLaunchKernelGGL(k_ntt_dit<false>, dim3(J.st[st].blocks), dim3(256), stage_lds(J, st),
                         stream, J, st);
    IMP_CHECK(hipGetLastError());
  }
  return hipSuccess;
}

// Scratch of one impeller_ntt call: the coset / inverse factor tables, n^-1
// and, for a fused-reversal plan, the 2^L-element second buffer.
static const size_t kFactorBytes = (size_t)3 * kFactorTab * 32;
static const size_t kNttFixedBytes = 2 * kFactorBytes + 256;
// the factor tables cover 3 * kFactorDigit = 39 index bits; transforms stop
// at 2^31 points (32-bit block counts), for every entry point
static const u32 kNoAllocMaxLog = 31;

struct NttScratch {
  u32 *pro, *epi, *ninv, *tmp;
};

static NttScratch carve_scratch(char* ws, size_t bytes, bool want_tmp, u32 L) {
  NttScratch s;
  s.pro = (u32*)ws;
  s.epi = (u32*)(ws + kFactorBytes);
  s.ninv = (u32*)(ws + 2 * kFactorBytes);
  s.tmp = nullptr;
  const size_t need = kNttFixedBytes + ((size_t)32 << L);
  if (want_tmp && bytes >= need) s.tmp = (u32*)(ws + kNttFixedBytes);
  return s;
}

// The factor tables of one transform go into scratch; returns the job ready
// to launch (its fused reversal dropped when the scratch has no room for it).
static hipError_t prepare_single(uint32_t* x, const uint32_t* twiddle, u32 L, int inverse,
                                 const uint32_t* coset_shift, const PassPlan& plan,
                                 const NttScratch& sc, hipStream_t stream, NttJob* out) {
  View v = {x, 0, 1ull << L};
  Twiddle tw = {twiddle, nullptr, nullptr, 0};
  Factor pro = {nullptr, 1, 1}, epi = {nullptr, 1, 1};
  if (!inverse) {
    if (coset_shift) {
      pro = factor_shape(sc.pro, L);
      IMP_CHECK(launch_factor_tables(pro, coset_shift, nullptr, stream));
    }
  } else {
    hipLaunchKernelGGL(k_ninv, dim3(1), dim3(1), 0, stream, sc.ninv, L);
    IMP_CHECK(hipGetLastError());
    // factor(i) = n^-1 * shift^i (shift given as its inverse by the caller)
    // or the constant n^-1 (base 1, one digit)
    epi = factor_shape(sc.epi, coset_shift ? L : 1);
    IMP_CHECK(launch_factor_tables(epi, coset_shift, sc.ninv, stream));
  }
  *out = make_job(v, v, v.wl, L, tw, false, pro, epi, plan);
  out->tmp = sc.tmp;
  if (out->fbr && !sc.tmp) drop_fused_reversal(out);
  return hipSuccess;
}

// Host Montgomery form of the integer x (< 2^64) for Fr
static host::Fe host_from_u64(uint64_t x) {
  host::Fe a = {{x, 0, 0, 0}}, r2;
  u32 t[8];
  for (int i = 0; i < 8; i++) t[i] = FrParams::r2[i];
  r2 = host::load(t);
  return host::mul<FrParams>(a, r2);
}

}  // namespace imp

using namespace imp;

// =============================================================== C ABI
extern "C" hipError_t ssip_ntt(PolyPtr x, const unsigned int* twiddle, unsigned int log_len,
                               hipStream_t stream, const unsigned int, const unsigned int) {
  IMP_REQUIRE(log_len <= 31, "log_len");
  IMP_REQUIRE(x.len == (1ull << log_len), "x.len must be 2^log_len");
  IMP_REQUIRE(x.ptr != nullptr && (log_len == 0 || twiddle != nullptr), "null pointer");
  View v;
  IMP_REQUIRE(make_view(x, &v), "PolyPtr view");
  Twiddle tw = {twiddle, nullptr, nullptr, 0};
  PassPlan plan;
  default_inplace_plan(log_len, &plan);
  Factor none = {nullptr, 1, 1};
  return launch_stages(make_job(v, v, x.len, log_len, tw, false, none, none, plan), stream);
}

extern "C" hipError_t recompute_ntt(PolyPtr x, const unsigned int* pq_d, unsigned int pq_deg,
                                    const unsigned int* omegas_d, unsigned int log_len,
                                    hipStream_t stream, const unsigned int, const unsigned int) {
  IMP_REQUIRE(log_len <= 31, "log_len");
  IMP_REQUIRE(x.len == (1ull << log_len), "x.len must be 2^log_len");
  IMP_REQUIRE(x.ptr != nullptr && omegas_d != nullptr, "null pointer");
  View v;
  IMP_REQUIRE(make_view(x, &v), "PolyPtr view");
  PassPlan plan;
  recompute_plan(log_len, &plan);
  Factor none = {nullptr, 1, 1};
  // pq only helps when it covers a pass's in-group twiddles (pq_deg >= D);
  // rc_prologue checks that per pass and otherwise uses omegas
  Twiddle tw = {nullptr, pq_deg >= 1 ? pq_d : nullptr, omegas_d, pq_deg};
  return launch_stages(make_job(v, v, x.len, log_len, tw, true, none, none, plan), stream);
}

extern "C" hipError_t ssip_precompute(unsigned int* twiddle, unsigned int len,
                                      const unsigned int* unit) {
  if (len == 0) return hipSuccess;
  IMP_REQUIRE(twiddle && unit, "null pointer");
  u32 *d_out = nullptr, *d_unit = nullptr;
  IMP_CHECK(hipMalloc((void**)&d_out, (size_t)len * 32));
  hipError_t e = hipMalloc((void**)&d_unit, 32);
  if (e == hipSuccess) e = hipMemcpy(d_unit, unit, 32, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    u32 chunk = 64;
    u64 threads = ((u64)len + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_powers, dim3((u32)((threads + 255) / 256)), dim3(256), 0, 0, d_out,
                       (u64)len, d_unit, chunk);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(twiddle, d_out, (size_t)len * 32, hipMemcpyDeviceToHost);
  (void)hipFree(d_out);
  if (d_unit) (void)hipFree(d_unit);
  return e;
}

// precompute.cuh:89-112 semantics, computed on the host: omegas[i] =
// unit^(2^i) for i < 32 (always), pq[0] = 1 and pq[i] = (unit^(len >>
// pq_deg))^i for 0 < i < 2^(pq_deg - 1).  Null pointers or pq_deg > 31 leave
// the outputs untouched (the function returns nothing to report with).
extern "C" void gen_pq_omegas(unsigned int* pq, unsigned int* omegas, unsigned int pq_deg,
                              unsigned int len, unsigned int* unit) {
  if (!pq || !omegas || !unit || pq_deg > 31) return;
  const host::Fe u = host::load(unit);
  host::Fe w = u;
  for (u32 i = 0; i < 32; i++) {
    host::store(omegas + 8 * i, w);
    w = host::mul<FrParams>(w, w);
  }
  host::store(pq, host::one<FrParams>());
  const u64 npq = (1ull << pq_deg) >> 1;
  if (npq < 2) return;
  const host::Fe step = host::pow<FrParams>(u, (u64)len >> pq_deg);
  host::Fe acc = step;
  for (u64 i = 1; i < npq; i++) {
    host::store(pq + 8 * i, acc);
    acc = host::mul<FrParams>(acc, step);
  }
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
  // omega^start for each thread's chunk comes from the device copy of omega,
  // placed past the table (the table's last chunk is written after the
  // kernel has read it: each thread reads omega before writing)
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

extern "C" size_t impeller_ntt_workspace_size(unsigned int log_len,
                                              const unsigned int* plan_passes,
                                              unsigned int n_passes) {
  PassPlan plan;
  if (log_len > 31 || !plan_from(plan_passes, n_passes, log_len, &plan)) return 0;
  return kNttFixedBytes + (plan.fuse == 1 ? ((size_t)32 << log_len) : 0);
}

extern "C" hipError_t impeller_ntt_ws(uint32_t* x, const uint32_t* twiddle, unsigned int log_len,
                                      int inverse, const uint32_t* coset_shift,
                                      const unsigned int* plan_passes, unsigned int n_passes,
                                      void* workspace, size_t workspace_bytes,
                                      hipStream_t stream) {
  IMP_REQUIRE(log_len <= 31, "log_len");
  IMP_REQUIRE(x && (log_len == 0 || twiddle), "null pointer");
  IMP_REQUIRE(workspace && workspace_bytes >= kNttFixedBytes, "workspace too small");
  PassPlan plan;
  IMP_REQUIRE(plan_from(plan_passes, n_passes, log_len, &plan), "pass plan");
  const NttScratch sc = carve_scratch((char*)workspace, workspace_bytes, plan.fuse == 1, log_len);
  NttJob J;
  IMP_CHECK(prepare_single(x, twiddle, log_len, inverse, coset_shift, plan, sc, stream, &J));
  return launch_stages(J, stream);
}

extern "C" hipError_t impeller_ntt(uint32_t* x, const uint32_t* twiddle, unsigned int log_len,
                                   int inverse, const uint32_t* coset_shift,
                                   const unsigned int* plan_passes, unsigned int n_passes,
                                   hipStream_t stream) {
  IMP_REQUIRE(log_len <= 31, "log_len");
  IMP_REQUIRE(x && (log_len == 0 || twiddle), "null pointer");
  PassPlan plan;
  IMP_REQUIRE(plan_from(plan_passes, n_passes, log_len, &plan), "pass plan");
  // scratch from the stream-ordered pool (its release threshold is left as
  // the process set it); without room for the second buffer the passes run
  // in place
  const bool need_factors = inverse || coset_shift;
  size_t bytes = plan.fuse == 1 ? kNttFixedBytes + ((size_t)32 << log_len) : kNttFixedBytes;
  if (!need_factors && plan.fuse != 1) bytes = 0;
  char* ws = nullptr;
  if (bytes && hipMallocAsync((void**)&ws, bytes, stream) != hipSuccess) {
    (void)hipGetLastError();
    ws = nullptr;
    if (need_factors) {
      bytes = kNttFixedBytes;
      IMP_CHECK(hipMallocAsync((void**)&ws, bytes, stream));
    }
  }
  NttScratch sc = {nullptr, nullptr, nullptr, nullptr};
  if (ws) sc = carve_scratch(ws, bytes, plan.fuse == 1, log_len);
  NttJob J;
  hipError_t e = prepare_single(x, twiddle, log_len, inverse, coset_shift, plan, sc, stream, &J);
  if (e == hipSuccess) e = launch_stages(J, stream);
  if (ws) {
    const hipError_t f = hipFreeAsync(ws, stream);
    if (e == hipSuccess) e = f;
  }
  return e;
}

// ------------------------------------------------- allocation-free variant
// impeller_ntt_noalloc keeps its scratch (the coset / inverse factor tables
// and n^-1: kNttFixedBytes = 1.5 MiB + 256 B) in this file-scope device
// array, one copy per GPU, held until the process ends.  Host code never
// passes the array's name to a kernel (on the host it only names the
// symbol): the device address comes from hipGetSymbolAddress, once per
// device, and is kept per device.
static const size_t kNoAllocWords = kNttFixedBytes / 4;
__device__ __align__(256) u32 scratch[kNoAllocWords];

namespace imp {
static const int kMaxDevices = 64;

static hipError_t noalloc_scratch(char** out) {
  static std::mutex mu;
  static char* ptr[kMaxDevices] = {};
  int dev = 0;
  IMP_CHECK(hipGetDevice(&dev));
  IMP_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal");
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    IMP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    IMP_REQUIRE(p != nullptr, "scratch symbol address");
    ptr[dev] = (char*)p;
  }
  *out = ptr[dev];
  return hipSuccess;
}
}  // namespace imp

extern "C" hipError_t impeller_ntt_noalloc(uint32_t* x, const uint32_t* twiddle,
                                           unsigned int log_len, int inverse,
                                           const uint32_t* coset_shift,
                                           const unsigned int* plan_passes,
                                           unsigned int n_passes, hipStream_t stream) {
  IMP_REQUIRE(log_len <= kNoAllocMaxLog, "log_len above the allocation-free maximum (31)");
  IMP_REQUIRE(x && (log_len == 0 || twiddle), "null pointer");
  PassPlan plan;
  IMP_REQUIRE(plan_from(plan_passes, n_passes, log_len, &plan), "pass plan");
  char* ws = nullptr;
  IMP_CHECK(noalloc_scratch(&ws));
  // the factor tables only (no room for a second buffer: a fused-reversal
  // plan runs its passes in place)
  const NttScratch sc = carve_scratch(ws, kNttFixedBytes, false, log_len);
  NttJob J;
  IMP_CHECK(prepare_single(x, twiddle, log_len, inverse, coset_shift, plan, sc, stream, &J));
  return launch_stages(J, stream);
}

// ------------------------------------------------------------ batched NTT
namespace {
const size_t kJobArea = kNttFixedBytes;  // pro table, epi table, n^-1
size_t batch_jobs_off() { return 0; }
size_t batch_prefix_off(u32 count) { return align_up((size_t)count * sizeof(NttJob), 256); }
size_t batch_factor_off(u32 count) {
  return batch_prefix_off(count) + align_up((size_t)9 * (count + 1) * 4, 256);
}
}  // namespace

extern "C" size_t impeller_ntt_batch_bytes(unsigned int count) {
  return batch_factor_off(count) + (size_t)count * kJobArea;
}

namespace {
// second buffer of an item whose plan fuses the bit reversal (0 otherwise)
size_t item_tmp_bytes(const ImpNttItem& it) {
  PassPlan plan;
  if (it.log_n > 31 || !plan_from(it.plan_len ? it.plan : nullptr, it.plan_len, it.log_n, &plan))
    return 0;
  return plan.fuse == 1 ? align_up((size_t)32 << it.log_n, 256) : 0;
}
}  // namespace

extern "C" size_t impeller_ntt_batch_bytes_items(const ImpNttItem* items, unsigned int count) {
  size_t b = align_up(impeller_ntt_batch_bytes(count), 256);
  for (u32 i = 0; items && i < count; i++) b += item_tmp_bytes(items[i]);
  return b;
}

extern "C" hipError_t impeller_ntt_batch_prepare(const ImpNttItem* items, unsigned int count,
                                                 void* dev, size_t bytes, ImpNttBatch* out,
                                                 hipStream_t stream) {
  IMP_REQUIRE(items && out && count > 0 && dev, "null pointer / empty batch");
  IMP_REQUIRE(bytes >= impeller_ntt_batch_bytes(count), "batch buffer too small");
  char* base = (char*)dev;
  std::vector<NttJob> jobs(count);
  std::vector<u32> prefix((size_t)9 * (count + 1), 0);
  u32 nst_max = 0;
  size_t tmp_off = align_up(impeller_ntt_batch_bytes(count), 256);
  for (u32 i = 0; i < count; i++) {
    const ImpNttItem& it = items[i];
    IMP_REQUIRE(it.data && it.log_n <= 31 && (it.log_n == 0 || it.twiddle), "batch item");
    IMP_REQUIRE(!it.src || it.log_src <= it.log_n, "batch item source size");
    PassPlan plan;
    IMP_REQUIRE(plan_from(it.plan_len ? it.plan : nullptr, it.plan_len, it.log_n, &plan),
                "batch item plan");
    // a fused-reversal plan takes its second buffer from the batch buffer
    // past the fixed area; without room it runs in place after the
    // bit-reversal stage (same passes, same result)
    u32* tmp = nullptr;
    if (plan.fuse == 1) {
      const size_t need = align_up((size_t)32 << it.log_n, 256);
      if (tmp_off + need <= bytes) {
        tmp = (u32*)(base + tmp_off);
        tmp_off += need;
      } else {
        plan.fuse = 0;
      }
    }
    View x = {it.data, 0, 1ull << it.log_n};
    View src = it.src ? View{const_cast<u32*>(it.src), 0, 1ull << it.log_src} : x;
    u64 src_len = it.src ? (1ull << it.log_src) : x.wl;
    Twiddle tw = {it.twiddle, nullptr, nullptr, 0};
    char* area = base + batch_factor_off(count) + (size_t)i * kJobArea;
    u32* pro_tab = (u32*)area;
    u32* epi_tab = (u32*)(area + kFactorBytes);
    u32* ninv = (u32*)(area + 2 * kFactorBytes);
    Factor pro = {nullptr, 1, 1}, epi = {nullptr, 1, 1};
    if (!it.inverse) {
      if (it.coset_shift) {
        pro = factor_shape(pro_tab, it.log_n);
        IMP_CHECK(launch_factor_tables(pro, it.coset_shift, nullptr, stream));
      }
    } else {
      hipLaunchKernelGGL(k_ninv, dim3(1), dim3(1), 0, stream, ninv, it.log_n);
      // factor(i) = n^-1 * shift^i, or the constant n^-1 (one 2-entry table)
      epi = factor_shape(epi_tab, it.coset_shift ? it.log_n : 1);
      IMP_CHECK(launch_factor_tables(epi, it.coset_shift, ninv, stream));
    }
    IMP_CHECK(hipGetLastError());
    jobs[i] = make_job(x, src, src_len, it.log_n, tw, false, pro, epi, plan);
    jobs[i].tmp = tmp;
    if (jobs[i].nst > nst_max) nst_max = jobs[i].nst;
  }
  out->dev = dev;
  out->count = count;
  out->nstages = 1 + nst_max;
  for (u32 st = 0; st < 9; st++) {
    u32 run = 0;
    size_t lds = 0;
    for (u32 i = 0; i < count; i++) {
      prefix[(size_t)st * (count + 1) + i] = run;
      u32 blocks = 0;
      if (st == 0) {
        if (!jobs[i].fbr) {
          blocks = jobs[i].br_blocks;
          size_t l = br_lds(jobs[i]);
          if (l > lds) lds = l;
        }
      } else if (st - 1 < jobs[i].nst) {
        blocks = jobs[i].st[st - 1].blocks;
        size_t l = stage_lds(jobs[i], st - 1);
        if (l > lds) lds = l;
      }
      run += blocks;
    }
    prefix[(size_t)st * (count + 1) + count] = run;
    out->total[st] = run;
    out->lds[st] = (unsigned int)lds;
  }
  IMP_CHECK(hipMemcpyAsync(base + batch_jobs_off(), jobs.data(), sizeof(NttJob) * count,
                           hipMemcpyHostToDevice, stream));
  IMP_CHECK(hipMemcpyAsync(base + batch_prefix_off(count), prefix.data(), prefix.size() * 4,
                           hipMemcpyHostToDevice, stream));
  // the host staging vectors die with this call
  return hipStreamSynchronize(stream);
}

extern "C" hipError_t impeller_ntt_batch_run(const ImpNttBatch* b, unsigned int max_blocks,
                                             hipStream_t stream) {
  IMP_REQUIRE(b && b->dev && b->count > 0, "batch");
  const char* base = (const char*)b->dev;
  const NttJob* jobs = (const NttJob*)(base + batch_jobs_off());
  const u32* prefix = (const u32*)(base + batch_prefix_off(b->count));
  for (u32 st = 0; st < b->nstages; st++) {
    u32 total = b->total[st];
    if (total == 0) continue;
    u32 grid = max_blocks && total > max_blocks ? max_blocks : total;
    const u32* pf = prefix + (size_t)st * (b->count + 1);
    if (st == 0) {
      TimedScope ts(K_NTT_BITREV, stream);
      hipLaunchKernelGGL(k_ntt_br_batch, dim3(grid), dim3(256), b->lds[st], stream, jobs,
                         b->count, pf);
    } else {
      TimedScope ts(K_NTT_DIT, stream);
      hipLaunchKernelGGL(k_ntt_dit_batch, dim3(grid), dim3(256), b->lds[st], stream, jobs,
                         b->count, pf, st - 1);
    }
    IMP_CHECK(hipGetLastError());
  }
  return hipSuccess;
}
