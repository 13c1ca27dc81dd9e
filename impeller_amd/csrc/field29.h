// field29.h -- BN254 Fr / Fq in 9 x 29-bit limbs, Montgomery form with
// R' = 2^261, for the multiply-bound inner loops (MSM point arithmetic).
//
// Why: with 32-bit limbs every partial product of a Montgomery multiply needs
// a carry instruction next to its v_mad_u64_u32.  With 29-bit limbs a whole
// product-scanning column (at most 9 a*b and 9 m*q products, each < 2^60, plus
// the previous column's carry) fits a 64-bit accumulator, so each partial
// product is one v_mad_u64_u32 and a column ends with one 64-bit shift.
//
// Values are lazy: limbs are kept below 2^29 ("normalized") except where a
// comment says otherwise, and the value may exceed the modulus m.  mul29
// returns a normalized value below 2m whenever a * b < m * R' (about 169 m^2
// for BN254), and accepts inputs whose limb bounds multiply to at most
// 1.5 * 2^60 (e.g. one normalized operand and one with limbs below 3 * 2^29).
// Subtraction adds a multiple of m whose limbs are padded so that no limb goes
// negative.  Values enter and leave through the 8 x 32-bit R = 2^256 form of
// the library's buffers: to_r29 multiplies by 2^266 mod m (R'^2 / R), from_r29
// by 2^256 mod m (R' R / R'), then reduces to canonical and repacks.
#pragma once
#include "field.h"

namespace imp {

static constexpr u32 kB29 = 29;
static constexpr u32 kM29 = (1u << 29) - 1;

struct Fr29Params {
  static constexpr u32 m[9] = {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u,
                               0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr u32 np = 0x0fffffffu;  // -m^-1 mod 2^29
  static constexpr u32 one[9] = {0x0fffff57u, 0x1ea70ab4u, 0x052c068bu, 0x17504f49u, 0x0aa8075bu,
                                 0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};
  static constexpr u32 kin[9] = {0x0fffead7u, 0x1d5444f4u, 0x04438aa5u, 0x03b4d096u, 0x134c84dau,
                                 0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};
  static constexpr u32 kout[9] = {0x0ffffffbu, 0x04b1a0e2u, 0x18334a6bu, 0x18ed2b3eu, 0x1462e36fu,
                                  0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};
  typedef FrParams P32;
};
struct Fq29Params {
  static constexpr u32 m[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                               0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr u32 np = 0x04866389u;
  static constexpr u32 one[9] = {0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u,
                                 0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};
  static constexpr u32 kin[9] = {0x13349ca1u, 0x1a5d84a8u, 0x0a3e5cacu, 0x100249e0u, 0x12b951e8u,
                                 0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};
  static constexpr u32 kout[9] = {0x058f0d9du, 0x1aea1c6eu, 0x11c2cf74u, 0x11d651ebu, 0x1462c0a7u,
                                  0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};
  typedef FqParams P32;
};

template <class P>
struct F29 {
  u32 v[9];
  __device__ __forceinline__ static F29 zero() {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = 0;
    return r;
  }
  __device__ __forceinline__ static F29 one() {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = P::one[i];
    return r;
  }
  __device__ __forceinline__ bool is_zero_exact() const {
    u32 acc = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) acc |= v[i];
    return acc == 0;
  }
};

typedef F29<Fr29Params> Fr29;
typedef F29<Fq29Params> Fq29;

// Montgomery product a b / 2^261 mod m (product scanning, 64-bit column
// accumulator, no carry instructions).  Column k < 9 adds m_k q_0 with
// m_k = acc q' mod 2^29, which clears its low 29 bits; columns 9 .. 16 give
// the result limbs, the final carry the top limb.
template <class P>
__device__ __forceinline__ F29<P> mul29(const F29<P>& a, const F29<P>& b) {
  u32 q[9];
  F29<P> r;
  u64 acc = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
#pragma unroll
    for (int j = 0; j < k; j++) {
      acc += (u64)a.v[j] * b.v[k - j];
      acc += (u64)q[j] * P::m[k - j];
    }
    acc += (u64)a.v[k] * b.v[0];
    q[k] = ((u32)acc * P::np) & kM29;
    acc += (u64)q[k] * P::m[0];
    acc >>= kB29;
  }
#pragma unroll
  for (int k = 9; k < 17; k++) {
#pragma unroll
    for (int j = k - 8; j < 9; j++) {
      acc += (u64)a.v[j] * b.v[k - j];
      acc += (u64)q[j] * P::m[k - j];
    }
    r.v[k - 9] = (u32)acc & kM29;
    acc >>= kB29;
  }
  r.v[8] = (u32)acc;
  return r;
}

// a^2: as mul29 with each cross product a_j a_(k-j), j < k - j, taken once
// and doubled (45 instead of 81 products for the a*a half).  a normalized.
template <class P>
__device__ __forceinline__ F29<P> sqr29(const F29<P>& a) {
  u32 q[9];
  F29<P> r;
  u64 acc = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    u64 off = 0;
#pragma unroll
    for (int j = 0; 2 * j < k; j++) off += (u64)a.v[j] * a.v[k - j];
    acc += off << 1;
    if ((k & 1) == 0) acc += (u64)a.v[k / 2] * a.v[k / 2];
#pragma unroll
    for (int j = 0; j < k; j++) acc += (u64)q[j] * P::m[k - j];
    q[k] = ((u32)acc * P::np) & kM29;
    acc += (u64)q[k] * P::m[0];
    acc >>= kB29;
  }
#pragma unroll
  for (int k = 9; k < 17; k++) {
    u64 off = 0;
#pragma unroll
    for (int j = k - 8; 2 * j < k; j++) off += (u64)a.v[j] * a.v[k - j];
    acc += off << 1;
    if ((k & 1) == 0) acc += (u64)a.v[k / 2] * a.v[k / 2];
#pragma unroll
    for (int j = k - 8; j < 9; j++) acc += (u64)q[j] * P::m[k - j];
    r.v[k - 9] = (u32)acc & kM29;
    acc >>= kB29;
  }
  r.v[8] = (u32)acc;
  return r;
}

// carries limbs 0..7 into the next limb (limbs below 2^32 on entry)
template <class P>
__device__ __forceinline__ void norm29(F29<P>& x) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    x.v[i + 1] += x.v[i] >> kB29;
    x.v[i] &= kM29;
  }
}

// limb-wise sum, not normalized (limbs below the sum of the inputs' bounds)
template <class P>
__device__ __forceinline__ F29<P> add29(const F29<P>& a, const F29<P>& b) {
  F29<P> r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.v[i] = a.v[i] + b.v[i];
  return r;
}

// a - b + K m, normalized.  K m is added with padded limbs (limb 0 + 2^29,
// limbs 1..7 + 2^29 - 1, limb 8 - 1: the same value) so that no limb below
// the top goes negative for a normalized b; a's limbs below 2^30; b < K m.
template <class P, int K>
__device__ __forceinline__ F29<P> sub29(const F29<P>& a, const F29<P>& b) {
  F29<P> r;
  u64 c = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    u64 t = (u64)P::m[i] * K + c;
    u32 km = i < 8 ? (u32)t & kM29 : (u32)t;
    c = t >> kB29;
    const u32 pad = i == 0 ? (1u << kB29) : (i < 8 ? (1u << kB29) - 1u : 0u);
    const u32 o = km + pad - (i == 8 ? 1u : 0u);
    r.v[i] = a.v[i] + o - b.v[i];
  }
  norm29<P>(r);
  return r;
}

// a - b - c - d + K m, normalized (padding 3 2^29 per limb; b, c, d
// normalized, b + c + d < K m)
template <class P, int K>
__device__ __forceinline__ F29<P> sub3_29(const F29<P>& a, const F29<P>& b, const F29<P>& c,
                                          const F29<P>& d) {
  F29<P> r;
  u64 cy = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    u64 t = (u64)P::m[i] * K + cy;
    u32 km = i < 8 ? (u32)t & kM29 : (u32)t;
    cy = t >> kB29;
    const u32 pad = i == 0 ? (3u << kB29) : (i < 8 ? (3u << kB29) - 3u : 0u);
    const u32 o = km + pad - (i == 8 ? 3u : 0u);
    r.v[i] = a.v[i] + o - b.v[i] - c.v[i] - d.v[i];
  }
  norm29<P>(r);
  return r;
}

// x -= K m when x >= K m (x normalized; signed borrows between limbs)
template <class P>
__device__ __forceinline__ void csub29(F29<P>& x, int K) {
  F29<P> d;
  int borrow = 0;
  u64 c = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    u64 t = (u64)P::m[i] * (u32)K + c;
    u32 km = i < 8 ? (u32)t & kM29 : (u32)t;
    c = t >> kB29;
    int s = (int)x.v[i] - (int)km + borrow;
    borrow = s >> kB29;  // arithmetic: -1 when negative
    d.v[i] = (u32)s & kM29;
    if (i == 8) borrow = s < 0 ? -1 : 0;
  }
  const bool keep = borrow == 0;
#pragma unroll
  for (int i = 0; i < 9; i++) x.v[i] = keep ? d.v[i] : x.v[i];
}

// canonical [0, m) from a normalized x < 8m
template <class P>
__device__ __forceinline__ F29<P> canon29(F29<P> x) {
  csub29<P>(x, 4);
  csub29<P>(x, 2);
  csub29<P>(x, 1);
  return x;
}

// x == 0 mod m for a normalized x < 2m
template <class P>
__device__ __forceinline__ bool is_zero29(const F29<P>& x) {
  u32 z = 0, e = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    z |= x.v[i];
    e |= x.v[i] ^ P::m[i];
  }
  return z == 0 || e == 0;
}

// 8 x 32 -> 9 x 29 bit fields (same integer)
template <class P>
__device__ __forceinline__ F29<P> unpack29(const u32 w[8]) {
  F29<P> r;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int bit = 29 * i, lo = bit >> 5, sh = bit & 31;
    u64 word = w[lo];
    if (lo + 1 < 8) word |= (u64)w[lo + 1] << 32;
    r.v[i] = (u32)(word >> sh) & kM29;
  }
  return r;
}

// 9 x 29 (normalized, value < 2^256) -> 8 x 32
template <class P>
__device__ __forceinline__ void pack29(const F29<P>& x, u32 w[8]) {
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int bit = 32 * j, lo = bit / 29, sh = bit % 29;
    u64 word = (u64)x.v[lo] >> sh;
    if (lo + 1 < 9) word |= (u64)x.v[lo + 1] << (29 - sh);
    if (lo + 2 < 9 && 58 - sh < 64) word |= (u64)x.v[lo + 2] << (58 - sh);
    w[j] = (u32)word;
  }
}

template <class P>
__device__ __forceinline__ F29<P> consts29(const u32* c) {
  F29<P> r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.v[i] = c[i];
  return r;
}

// R-form 8 x 32 element (any representative below 2^256) -> R'-form, < 2m
template <class P>
__device__ __forceinline__ F29<P> to_r29(const Fp<typename P::P32>& x) {
  return mul29<P>(unpack29<P>(x.v), consts29<P>(P::kin));
}

// R'-form (normalized, < 8m) -> canonical R-form 8 x 32
template <class P>
__device__ __forceinline__ Fp<typename P::P32> from_r29(const F29<P>& x) {
  F29<P> y = canon29<P>(mul29<P>(x, consts29<P>(P::kout)));
  Fp<typename P::P32> r;
  pack29<P>(y, r.v);
  return r;
}

// ------------------------------------------------------------ G1, 29-bit
// The XYZZ formulas of field.h (madd-2008-s, add-2008-s, dbl-2008-s-1,
// mdbl-2008-s-1) over Fq29, with every coordinate normalized and
//   x < 8m, y < 4m, zz < 2m, zzz < 2m      (XYZZ values),
//   x, y < 2m                               (affine values).
// Each product below is annotated with the bound that keeps it under
// m R' (output < 2m).  The identity is zz == 0 exactly (XYZZ) and
// x == y == 0 exactly (affine); every coordinate is the same field element as
// the 32-bit path computes, so results agree after from_r29.
struct G1Aff29 {
  Fq29 x, y;
  __device__ __forceinline__ bool is_identity() const {
    return x.is_zero_exact() && y.is_zero_exact();
  }
};
struct G1Xyzz29 {
  Fq29 x, y, zz, zzz;
  __device__ __forceinline__ static G1Xyzz29 identity() {
    G1Xyzz29 r;
    r.x = Fq29::zero();
    r.y = Fq29::zero();
    r.zz = Fq29::zero();
    r.zzz = Fq29::one();
    return r;
  }
  __device__ __forceinline__ bool is_identity() const { return zz.is_zero_exact(); }
};

typedef Fq29Params Q29;

__device__ __forceinline__ G1Xyzz29 aff29_to_xyzz(const G1Aff29& a) {
  if (a.is_identity()) return G1Xyzz29::identity();
  G1Xyzz29 r;
  r.x = a.x;
  r.y = a.y;
  r.zz = Fq29::one();
  r.zzz = Fq29::one();
  return r;
}

// -y for an affine point (y < m canonical, or the identity)
__device__ __forceinline__ G1Aff29 aff29_neg(const G1Aff29& a) {
  if (a.is_identity()) return a;
  G1Aff29 r;
  r.x = a.x;
  r.y = sub29<Q29, 1>(Fq29::zero(), a.y);
  return r;
}

// mdbl-2008-s-1: 2 * affine point (x, y < 2m)
__device__ __forceinline__ G1Xyzz29 aff29_dbl(const G1Aff29& p) {
  const Fq29 zero = Fq29::zero();
  Fq29 u = add29<Q29>(p.y, p.y);                   // limbs < 2^30, < 4m
  Fq29 v = mul29<Q29>(u, u);                       // 16 m^2
  Fq29 w = mul29<Q29>(u, v);                       // 8 m^2
  Fq29 s = mul29<Q29>(p.x, v);                     // 4 m^2
  Fq29 x2 = mul29<Q29>(p.x, p.x);                  // 4 m^2
  Fq29 m3 = add29<Q29>(add29<Q29>(x2, x2), x2);
  norm29<Q29>(m3);                                 // < 6m
  Fq29 x3 = sub3_29<Q29, 4>(mul29<Q29>(m3, m3), s, s, zero);      // 36 m^2; < 6m
  Fq29 t = sub29<Q29, 8>(s, x3);                                   // < 10m
  Fq29 y3 = sub29<Q29, 2>(mul29<Q29>(m3, t), mul29<Q29>(w, p.y));  // 60, 4 m^2; < 4m
  G1Xyzz29 r;
  r.x = x3;
  r.y = y3;
  r.zz = v;
  r.zzz = w;
  return r;
}

// dbl-2008-s-1
__device__ __forceinline__ G1Xyzz29 xyzz29_dbl(const G1Xyzz29& p) {
  if (p.is_identity()) return p;
  const Fq29 zero = Fq29::zero();
  Fq29 u = add29<Q29>(p.y, p.y);                   // limbs < 2^30, < 8m
  Fq29 v = mul29<Q29>(u, u);                       // 64 m^2
  Fq29 w = mul29<Q29>(u, v);                       // 16 m^2
  Fq29 s = mul29<Q29>(p.x, v);                     // 16 m^2
  Fq29 x2 = sqr29<Q29>(p.x);                       // 64 m^2
  Fq29 m3 = add29<Q29>(add29<Q29>(x2, x2), x2);
  norm29<Q29>(m3);                                 // < 6m
  Fq29 x3 = sub3_29<Q29, 4>(sqr29<Q29>(m3), s, s, zero);          // 36 m^2; < 6m
  Fq29 t = sub29<Q29, 8>(s, x3);                                   // < 10m
  Fq29 y3 = sub29<Q29, 2>(mul29<Q29>(m3, t), mul29<Q29>(w, p.y));  // 60, 8 m^2; < 4m
  G1Xyzz29 r;
  r.x = x3;
  r.y = y3;
  r.zz = mul29<Q29>(v, p.zz);                      // 4 m^2
  r.zzz = mul29<Q29>(w, p.zzz);                    // 4 m^2
  return r;
}

// madd-2008-s: XYZZ + affine (q: x, y < 2m)
__device__ __forceinline__ G1Xyzz29 xyzz29_madd(const G1Xyzz29& p, const G1Aff29& q) {
  if (q.is_identity()) return p;
  if (p.is_identity()) return aff29_to_xyzz(q);
  Fq29 u2 = mul29<Q29>(q.x, p.zz);                 // 4 m^2
  Fq29 s2 = mul29<Q29>(q.y, p.zzz);                // 4 m^2
  Fq29 P_ = sub29<Q29, 8>(u2, p.x);                // < 10m
  Fq29 R_ = sub29<Q29, 4>(s2, p.y);                // < 6m
  Fq29 pp = sqr29<Q29>(P_);                        // 100 m^2
  Fq29 rr = sqr29<Q29>(R_);                        // 36 m^2
  if (is_zero29<Q29>(pp)) {
    if (is_zero29<Q29>(rr)) return aff29_dbl(q);
    return G1Xyzz29::identity();
  }
  Fq29 ppp = mul29<Q29>(P_, pp);                   // 20 m^2
  Fq29 qq = mul29<Q29>(p.x, pp);                   // 16 m^2
  Fq29 x3 = sub3_29<Q29, 6>(rr, ppp, qq, qq);      // < 8m
  Fq29 t = sub29<Q29, 8>(qq, x3);                  // < 10m
  Fq29 y3 = sub29<Q29, 2>(mul29<Q29>(R_, t), mul29<Q29>(p.y, ppp));  // 60, 8 m^2; < 4m
  G1Xyzz29 r;
  r.x = x3;
  r.y = y3;
  r.zz = mul29<Q29>(p.zz, pp);                     // 4 m^2
  r.zzz = mul29<Q29>(p.zzz, ppp);                  // 4 m^2
  return r;
}

// add-2008-s: XYZZ + XYZZ
__device__ __forceinline__ G1Xyzz29 xyzz29_add(const G1Xyzz29& p, const G1Xyzz29& q) {
  if (p.is_identity()) return q;
  if (q.is_identity()) return p;
  Fq29 u1 = mul29<Q29>(p.x, q.zz);                 // 16 m^2
  Fq29 u2 = mul29<Q29>(q.x, p.zz);                 // 16 m^2
  Fq29 s1 = mul29<Q29>(p.y, q.zzz);                // 8 m^2
  Fq29 s2 = mul29<Q29>(q.y, p.zzz);                // 8 m^2
  Fq29 P_ = sub29<Q29, 2>(u2, u1);                 // < 4m
  Fq29 R_ = sub29<Q29, 2>(s2, s1);                 // < 4m
  Fq29 pp = sqr29<Q29>(P_);                        // 16 m^2
  Fq29 rr = sqr29<Q29>(R_);                        // 16 m^2
  if (is_zero29<Q29>(pp)) {
    if (is_zero29<Q29>(rr)) return xyzz29_dbl(p);
    return G1Xyzz29::identity();
  }
  Fq29 ppp = mul29<Q29>(P_, pp);                   // 8 m^2
  Fq29 qq = mul29<Q29>(u1, pp);                    // 4 m^2
  Fq29 x3 = sub3_29<Q29, 6>(rr, ppp, qq, qq);      // < 8m
  Fq29 t = sub29<Q29, 8>(qq, x3);                  // < 10m
  Fq29 y3 = sub29<Q29, 2>(mul29<Q29>(R_, t), mul29<Q29>(s1, ppp));  // 40, 4 m^2; < 4m
  G1Xyzz29 r;
  r.x = x3;
  r.y = y3;
  r.zz = mul29<Q29>(mul29<Q29>(p.zz, q.zz), pp);   // 4, 4 m^2
  r.zzz = mul29<Q29>(mul29<Q29>(p.zzz, q.zzz), ppp);
  return r;
}

// ------------------------------------------- group-of-four (quad) forms
// xyzz29_dbl / xyzz29_add with their products split into rounds of at most
// four independent ones, one per lane of a group of four lanes that all hold
// the same points (q = lane % 4), the round's results broadcast to the whole
// group (quad_bcast, field.h).  For chains of dependent operations on few
// lanes, where one wave's latency sets the time.  Every product has the
// operands of the corresponding product in xyzz29_dbl / xyzz29_add (a
// square is taken as the product of equal operands), so the bounds noted
// there hold; the values are the same field elements.
template <int K>
__device__ __forceinline__ Fq29 quad_fe29(const Fq29& a) {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.v[i] = quad_bcast<K>(a.v[i]);
  return r;
}
// a_q for q = lane % 4, as two levels of selects on the bits of q (one or
// two v_cndmask per limb); the form sel(q, a0, a1, a2, a3) written as a
// chain of comparisons with q is turned by the compiler into an index into
// a private array (scratch memory)
__device__ __forceinline__ Fq29 pick4_29(u32 q, const Fq29& a0, const Fq29& a1, const Fq29& a2,
                                         const Fq29& a3) {
  const bool b0 = (q & 1) != 0, b1 = (q & 2) != 0;
  Fq29 r;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const u32 lo = b0 ? a1.v[i] : a0.v[i], hi = b0 ? a3.v[i] : a2.v[i];
    r.v[i] = b1 ? hi : lo;
  }
  return r;
}

// dbl-2008-s-1: {u u, x x}, {u v, x v, v zz, m3 m3}, {m3 t, w y, w zzz}
__device__ __forceinline__ G1Xyzz29 xyzz29_dbl_quad(const G1Xyzz29& p, u32 q) {
  if (p.is_identity()) return p;
  const Fq29 zero = Fq29::zero();
  const Fq29 u = add29<Q29>(p.y, p.y);
  const Fq29 a1 = pick4_29(q, u, p.x, u, p.x);
  const Fq29 r1 = mul29<Q29>(a1, a1);
  const Fq29 v = quad_fe29<0>(r1), x2 = quad_fe29<1>(r1);
  Fq29 m3 = add29<Q29>(add29<Q29>(x2, x2), x2);
  norm29<Q29>(m3);
  const Fq29 r2 = mul29<Q29>(pick4_29(q, u, p.x, v, m3), pick4_29(q, v, v, p.zz, m3));
  const Fq29 w = quad_fe29<0>(r2), s = quad_fe29<1>(r2), zz3 = quad_fe29<2>(r2),
             msq = quad_fe29<3>(r2);
  const Fq29 x3 = sub3_29<Q29, 4>(msq, s, s, zero);
  const Fq29 t = sub29<Q29, 8>(s, x3);
  const Fq29 r3 = mul29<Q29>(pick4_29(q, m3, w, w, w), pick4_29(q, t, p.y, p.zzz, p.zzz));
  G1Xyzz29 r;
  r.x = x3;
  r.y = sub29<Q29, 2>(quad_fe29<0>(r3), quad_fe29<1>(r3));
  r.zz = zz3;
  r.zzz = quad_fe29<2>(r3);
  return r;
}

// add-2008-s: {u1, u2, s1, s2}, {P P, R R, zz1 zz2, zzz1 zzz2},
// {P pp, u1 pp, zz12 pp}, {R t, s1 ppp, zzz12 ppp}
__device__ __forceinline__ G1Xyzz29 xyzz29_add_quad(const G1Xyzz29& p, const G1Xyzz29& o, u32 q) {
  if (p.is_identity()) return o;
  if (o.is_identity()) return p;
  const Fq29 r1 = mul29<Q29>(pick4_29(q, p.x, o.x, p.y, o.y), pick4_29(q, o.zz, p.zz, o.zzz, p.zzz));
  const Fq29 u1 = quad_fe29<0>(r1), u2 = quad_fe29<1>(r1), s1 = quad_fe29<2>(r1),
             s2 = quad_fe29<3>(r1);
  const Fq29 P_ = sub29<Q29, 2>(u2, u1);
  const Fq29 R_ = sub29<Q29, 2>(s2, s1);
  const Fq29 r2 = mul29<Q29>(pick4_29(q, P_, R_, p.zz, p.zzz), pick4_29(q, P_, R_, o.zz, o.zzz));
  const Fq29 pp = quad_fe29<0>(r2), rr = quad_fe29<1>(r2), zz12 = quad_fe29<2>(r2),
             zzz12 = quad_fe29<3>(r2);
  if (is_zero29<Q29>(pp)) {
    if (is_zero29<Q29>(rr)) return xyzz29_dbl_quad(p, q);
    return G1Xyzz29::identity();
  }
  const Fq29 r3 = mul29<Q29>(pick4_29(q, P_, u1, zz12, zz12), pp);
  const Fq29 ppp = quad_fe29<0>(r3), qq = quad_fe29<1>(r3), zz3 = quad_fe29<2>(r3);
  const Fq29 x3 = sub3_29<Q29, 6>(rr, ppp, qq, qq);
  const Fq29 t = sub29<Q29, 8>(qq, x3);
  const Fq29 r4 = mul29<Q29>(pick4_29(q, R_, s1, zzz12, zzz12), pick4_29(q, t, ppp, ppp, ppp));
  G1Xyzz29 r;
  r.x = x3;
  r.y = sub29<Q29, 2>(quad_fe29<0>(r4), quad_fe29<1>(r4));
  r.zz = zz3;
  r.zzz = quad_fe29<2>(r4);
  return r;
}

// ---------------------------------------------------- conversions / memory
// XYZZ29 in memory: 4 x 9 u32 (144 bytes), limbs as held
__device__ __forceinline__ void store_xyzz29(u32* p, const G1Xyzz29& a) {
#pragma unroll
  for (int i = 0; i < 9; i++) {
    p[i] = a.x.v[i];
    p[9 + i] = a.y.v[i];
    p[18 + i] = a.zz.v[i];
    p[27 + i] = a.zzz.v[i];
  }
}
// the same layout in global memory (p 16-byte aligned): nine 16-byte
// accesses through a global-space pointer
__device__ __forceinline__ void store_xyzz29_g(u32* p, const G1Xyzz29& a) {
  u32 w[36];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    w[i] = a.x.v[i];
    w[9 + i] = a.y.v[i];
    w[18 + i] = a.zz.v[i];
    w[27 + i] = a.zzz.v[i];
  }
  uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
  for (int k = 0; k < 9; k++) gst(q + k, make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]));
}
__device__ __forceinline__ G1Xyzz29 load_xyzz29_g(const u32* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  u32 w[36];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const uint4 v = gld(q + k);
    w[4 * k] = v.x;
    w[4 * k + 1] = v.y;
    w[4 * k + 2] = v.z;
    w[4 * k + 3] = v.w;
  }
  G1Xyzz29 a;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    a.x.v[i] = w[i];
    a.y.v[i] = w[9 + i];
    a.zz.v[i] = w[18 + i];
    a.zzz.v[i] = w[27 + i];
  }
  return a;
}
__device__ __forceinline__ G1Xyzz29 load_xyzz29(const u32* p) {
  G1Xyzz29 a;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    a.x.v[i] = p[i];
    a.y.v[i] = p[9 + i];
    a.zz.v[i] = p[18 + i];
    a.zzz.v[i] = p[27 + i];
  }
  return a;
}

// table format: affine point, each coordinate canonical in R'-form, stored
// as the 256-bit integer in 8 x u32 (64 bytes, like the R-form input)
__device__ __forceinline__ G1Aff29 load_aff_table(const u32* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 a = q[0], b = q[1], c = q[2], d = q[3];
  const u32 x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  const u32 y[8] = {c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w};
  G1Aff29 r;
  r.x = unpack29<Q29>(x);
  r.y = unpack29<Q29>(y);
  return r;
}

// R-form affine point -> table format (identity (0, 0) stays (0, 0))
__device__ __forceinline__ void aff_to_table(u32* out, const G1Aff& a) {
  u32 w[8];
  pack29<Q29>(canon29<Q29>(to_r29<Q29>(a.x)), w);
  store_fe<Fq>(out, *reinterpret_cast<const Fq*>(w));
  pack29<Q29>(canon29<Q29>(to_r29<Q29>(a.y)), w);
  store_fe<Fq>(out + 8, *reinterpret_cast<const Fq*>(w));
}

__device__ __forceinline__ G1Xyzz29 xyzz_to_29(const G1Xyzz& a) {
  G1Xyzz29 r;
  r.x = to_r29<Q29>(a.x);
  r.y = to_r29<Q29>(a.y);
  r.zz = to_r29<Q29>(a.zz);
  r.zzz = to_r29<Q29>(a.zzz);
  return r;
}

__device__ __forceinline__ G1Xyzz xyzz_from_29(const G1Xyzz29& a) {
  G1Xyzz r;
  r.x = from_r29<Q29>(a.x);
  r.y = from_r29<Q29>(a.y);
  r.zz = from_r29<Q29>(a.zz);
  r.zzz = from_r29<Q29>(a.zzz);
  return r;
}

}  // namespace imp
