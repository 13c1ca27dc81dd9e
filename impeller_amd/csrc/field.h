// field.h -- BN254 Fr / Fq Montgomery arithmetic and G1 XYZZ group law for
// gfx950 (wave64 VALU, 8 x u32 limbs).
//
// Representation matches the reference (core/src/common/mont/src/field.cuh:
// 638-743): 8 little-endian u32 limbs, Montgomery form with R = 2^256,
// canonical (< modulus) at every API boundary.  Multiplication is the CIOS
// Montgomery product (Koc, Acar, Kaliski 1996) in its "no final carry word"
// form, which is valid because the top limb of both BN254 moduli is below
// 2^31 - 1; the result of one product is < 2m and one conditional subtraction
// makes it canonical.  No MFMA: this is exact big-integer work.
//
// Curve formulas are the EFD short-Weierstrass XYZZ set the reference uses
// (core/src/common/curve/src/curve_xyzz.cuh:233-391): madd-2008-s,
// add-2008-s, dbl-2008-s-1, mdbl-2008-s-1, affine conversion by z-scaling.
// Identity: XYZZ with zz == 0; affine (0, 0).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace imp {

typedef uint32_t u32;
typedef uint64_t u64;

// Constants: core/src/common/mont/src/bn254_fr.cuh:21-25,67 and
// bn254_fq.cuh:13-17,59 (little-endian limbs).
struct FrParams {
  static constexpr u32 m[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                               0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr u32 np = 4026531839u;  // -m^-1 mod 2^32
  static constexpr u32 r1[8] = {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
                                0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
  static constexpr u32 r2[8] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                                0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};
};
struct FqParams {
  static constexpr u32 m[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                               0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr u32 np = 3834012553u;
  static constexpr u32 r1[8] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
  static constexpr u32 r2[8] = {0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
                                0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u};
};

template <class P>
struct Fp {
  u32 v[8];

  __host__ __device__ __forceinline__ static Fp zero() {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = 0;
    return r;
  }
  __host__ __device__ __forceinline__ static Fp one() {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = P::r1[i];
    return r;
  }
  __host__ __device__ __forceinline__ static Fp r2() {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = P::r2[i];
    return r;
  }
  __device__ __forceinline__ bool is_zero() const {
    u32 acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) acc |= v[i];
    return acc == 0;
  }
  __device__ __forceinline__ bool operator==(const Fp& o) const {
    u32 acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) acc |= v[i] ^ o.v[i];
    return acc == 0;
  }
};

typedef Fp<FrParams> Fr;
typedef Fp<FqParams> Fq;

// ---------------------------------------------------------------- loads/stores
// 32-byte element as two 16-byte vector accesses.
template <class F>
__device__ __forceinline__ F load_fe(const u32* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 a = q[0], b = q[1];
  F r;
  r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
  r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
  return r;
}
template <class F>
__device__ __forceinline__ void store_fe(u32* p, const F& x) {
  uint4* q = reinterpret_cast<uint4*>(p);
  q[0] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  q[1] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}
// Non-temporal variants for data touched once per pass.
template <class F>
__device__ __forceinline__ F load_fe_nt(const u32* p) {
  F r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.v[i] = __builtin_nontemporal_load(p + i);
  return r;
}

// --------------------------------------------------------------- arithmetic
// 32-bit add (subtract) with carry (borrow) in and out, carries 0 or 1.  A
// chain of these compiles to one v_add_co / v_addc_co (v_sub_co / v_subb_co)
// per limb; the same chain written with 64-bit intermediates costs 64-bit
// adds, shifts and moves per limb.
__host__ __device__ __forceinline__ u32 adc(u32 a, u32 b, u32 cin, u32* cout) {
  return __builtin_addc(a, b, cin, cout);
}
__host__ __device__ __forceinline__ u32 sbb(u32 a, u32 b, u32 bin, u32* bout) {
  return __builtin_subc(a, b, bin, bout);
}

// r = a - m if a >= m else a   (a < 2m)
template <class P>
__device__ __forceinline__ Fp<P> reduce_once(const Fp<P>& a) {
  Fp<P> d;
  u32 borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) d.v[i] = sbb(a.v[i], P::m[i], borrow, &borrow);
  Fp<P> r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.v[i] = borrow ? a.v[i] : d.v[i];
  return r;
}

template <class P>
__device__ __forceinline__ Fp<P> add(const Fp<P>& a, const Fp<P>& b) {
  // a, b < m < 2^254 so a + b < 2^255: no carry out of the top limb.
  Fp<P> s;
  u32 c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) s.v[i] = adc(a.v[i], b.v[i], c, &c);
  return reduce_once<P>(s);
}

template <class P>
__device__ __forceinline__ Fp<P> sub(const Fp<P>& a, const Fp<P>& b) {
  Fp<P> d;
  u32 borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) d.v[i] = sbb(a.v[i], b.v[i], borrow, &borrow);
  // add back m when the subtraction wrapped
  const u32 mask = 0u - borrow;
  Fp<P> r;
  u32 c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.v[i] = adc(d.v[i], P::m[i] & mask, c, &c);
  return r;
}

template <class P>
__device__ __forceinline__ Fp<P> neg(const Fp<P>& a) {
  if (a.is_zero()) return a;
  Fp<P> r;
  u32 borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.v[i] = sbb(P::m[i], a.v[i], borrow, &borrow);
  return r;
}

template <class P>
__device__ __forceinline__ Fp<P> dbl(const Fp<P>& a) {
  return add<P>(a, a);
}

// Montgomery product a*b*2^-256 mod m, result in [0, 2m) (no final subtract).
//
// Operand-scanning CIOS in plain C (kept as the portable reference form).
template <class P>
__device__ __forceinline__ Fp<P> mul_lazy_cios(const Fp<P>& a, const Fp<P>& b) {
  u32 t[8];
#pragma unroll
  for (int j = 0; j < 8; j++) t[j] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    u64 A = (u64)a.v[0] * b.v[i] + t[0];
    t[0] = (u32)A;
    A >>= 32;
    u32 q = t[0] * P::np;
    u64 C = ((u64)q * P::m[0] + t[0]) >> 32;
#pragma unroll
    for (int j = 1; j < 8; j++) {
      A = (u64)a.v[j] * b.v[i] + t[j] + A;
      t[j] = (u32)A;
      A >>= 32;
      C = (u64)q * P::m[j] + t[j] + C;
      t[j - 1] = (u32)C;
      C >>= 32;
    }
    t[7] = (u32)(C + A);
  }
  Fp<P> r;
#pragma unroll
  for (int j = 0; j < 8; j++) r.v[j] = t[j];
  return r;
}

#ifndef IMP_EMU
// (acc, hi) += x * y with a 96-bit accumulator: one v_mad_u64_u32 whose
// carry-out feeds one v_addc_co_u32.  The 64-bit accumulator is the mad's own
// addend register pair, so no zero-extension moves are needed.  gfx950 needs
// two wait states between a VALU write of an SGPR (the carry-out) and a VALU
// read of it (the carry-in); the compiler's hazard recognizer does not look
// inside an asm block, so the s_nop is written here.
__device__ __forceinline__ void mac96(u64& acc, u32& hi, u32 x, u32 y) {
  u64 c;
  asm("v_mad_u64_u32 %0, %1, %3, %4, %0\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %2, %1, %2, 0, %1"
      : "+v"(acc), "=&s"(c), "+v"(hi)
      : "v"(x), "v"(y));
}
// same with a wave-uniform (SGPR) second factor, e.g. a modulus limb
__device__ __forceinline__ void mac96s(u64& acc, u32& hi, u32 x, u32 y) {
  u64 c;
  asm("v_mad_u64_u32 %0, %1, %3, %4, %0\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %2, %1, %2, 0, %1"
      : "+v"(acc), "=&s"(c), "+v"(hi)
      : "v"(x), "s"(y));
}

// Two and four products per asm block, each mad with its own carry SGPR pair,
// ordered so that every carry read comes at least two wait states after its
// mad: one s_nop for the pair, none for the quad.  Factors: x*y with y in a
// VGPR, then u*w with w wave-uniform (a modulus limb).
__device__ __forceinline__ void mac96_pair(u64& acc, u32& hi, u32 x, u32 y, u32 u, u32 w) {
  u64 c0, c1;
  asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\t"
      "v_mad_u64_u32 %0, %3, %6, %7, %0\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32 %1, %2, %1, 0, %2\n\t"
      "v_addc_co_u32 %1, %3, %1, 0, %3"
      : "+v"(acc), "+v"(hi), "=&s"(c0), "=&s"(c1)
      : "v"(x), "v"(y), "v"(u), "s"(w));
}
__device__ __forceinline__ void mac96_quad(u64& acc, u32& hi, u32 x0, u32 y0, u32 u0, u32 w0,
                                           u32 x1, u32 y1, u32 u1, u32 w1) {
  u64 c0, c1, c2, c3;
  asm("v_mad_u64_u32 %0, %2, %6, %7, %0\n\t"
      "v_mad_u64_u32 %0, %3, %8, %9, %0\n\t"
      "v_mad_u64_u32 %0, %4, %10, %11, %0\n\t"
      "v_mad_u64_u32 %0, %5, %12, %13, %0\n\t"
      "v_addc_co_u32 %1, %2, %1, 0, %2\n\t"
      "v_addc_co_u32 %1, %3, %1, 0, %3\n\t"
      "v_addc_co_u32 %1, %4, %1, 0, %4\n\t"
      "v_addc_co_u32 %1, %5, %1, 0, %5"
      : "+v"(acc), "+v"(hi), "=&s"(c0), "=&s"(c1), "=&s"(c2), "=&s"(c3)
      : "v"(x0), "v"(y0), "v"(u0), "s"(w0), "v"(x1), "v"(y1), "v"(u1), "s"(w1));
}

// Product-scanning Montgomery multiplication (FIPS, Koc-Acar-Kaliski 1996):
// column k accumulates a_j*b_(k-j) and m_j*q_(k-j) into a 96-bit accumulator;
// m_k = acc * q' mod 2^32 for k < 8.  The result is (a b + q m) / 2^256 <
// a b / 2^256 + m: below 2m for a, b < 2m and 4m < 2^256, and also for a < 4m,
// b < m (the NTT's lazy butterflies); it fits 8 limbs.  The (a_j b, m_j q) pairs of a column go through
// the quad and pair blocks above; the column's last products (a_k b_0, then
// m_k q_0 once m_k is known) through the single-product forms.
template <class P>
__device__ __forceinline__ Fp<P> mul_lazy_fips(const Fp<P>& a, const Fp<P>& b) {
  u32 m[8];
  Fp<P> r;
  u64 acc = 0;
  u32 hi = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
#pragma unroll
    for (int j = 0; j < i; j += 2) {
      if (j + 1 < i)
        mac96_quad(acc, hi, a.v[j], b.v[i - j], m[j], P::m[i - j], a.v[j + 1], b.v[i - j - 1],
                   m[j + 1], P::m[i - j - 1]);
      else
        mac96_pair(acc, hi, a.v[j], b.v[i - j], m[j], P::m[i - j]);
    }
    mac96(acc, hi, a.v[i], b.v[0]);
    m[i] = (u32)acc * P::np;
    mac96s(acc, hi, m[i], P::m[0]);
    acc = (acc >> 32) | ((u64)hi << 32);
    hi = 0;
  }
#pragma unroll
  for (int i = 8; i < 16; i++) {
#pragma unroll
    for (int j = i - 7; j < 8; j += 2) {
      if (j + 1 < 8)
        mac96_quad(acc, hi, a.v[j], b.v[i - j], m[j], P::m[i - j], a.v[j + 1], b.v[i - j - 1],
                   m[j + 1], P::m[i - j - 1]);
      else
        mac96_pair(acc, hi, a.v[j], b.v[i - j], m[j], P::m[i - j]);
    }
    r.v[i - 8] = (u32)acc;
    acc = (acc >> 32) | ((u64)hi << 32);
    hi = 0;
  }
  return r;
}

#endif  // !IMP_EMU

// The library uses the inline-asm product-scanning form (1.43x the
// compiler-scheduled CIOS form on gfx950, bit-identical).  The CIOS form is
// kept for the host build of the kernels (IMP_EMU: the CPU launch emulation
// used by the tests) and as a measurement reference (IMP_MONT_CIOS).
template <class P>
__device__ __forceinline__ Fp<P> mul_lazy(const Fp<P>& a, const Fp<P>& b) {
#if defined(IMP_EMU) || defined(IMP_MONT_CIOS)
  return mul_lazy_cios<P>(a, b);
#else
  return mul_lazy_fips<P>(a, b);
#endif
}

template <class P>
__device__ __forceinline__ Fp<P> mul(const Fp<P>& a, const Fp<P>& b) {
  return reduce_once<P>(mul_lazy<P>(a, b));
}
template <class P>
__device__ __forceinline__ Fp<P> sqr(const Fp<P>& a) {
  return mul<P>(a, a);
}

// Montgomery -> plain integer (mul by 1), plain integer -> Montgomery (mul by R^2)
template <class P>
__device__ __forceinline__ Fp<P> to_number(const Fp<P>& a) {
  Fp<P> one_int = Fp<P>::zero();
  one_int.v[0] = 1;
  return mul<P>(a, one_int);
}
template <class P>
__device__ __forceinline__ Fp<P> from_number(const Fp<P>& a) {
  return mul<P>(a, Fp<P>::r2());
}

// a^e for a 32-bit exponent, square-and-multiply from the top bit
template <class P>
__device__ __forceinline__ Fp<P> pow_u32(const Fp<P>& a, u32 e) {
  Fp<P> acc = Fp<P>::one();
  for (int i = 31; i >= 0; i--) {
    acc = sqr<P>(acc);
    if ((e >> i) & 1) acc = mul<P>(acc, a);
  }
  return acc;
}

// a^e for a 256-bit exponent given as limbs
template <class P>
__device__ __forceinline__ Fp<P> pow_big(const Fp<P>& a, const u32 e[8]) {
  Fp<P> acc = Fp<P>::one();
  for (int w = 7; w >= 0; w--) {
    for (int i = 31; i >= 0; i--) {
      acc = sqr<P>(acc);
      if ((e[w] >> i) & 1) acc = mul<P>(acc, a);
    }
  }
  return acc;
}

// inversion by Fermat: a^(m-2) (field.cuh:1172-1175); inv(0) = 0.  Kept as
// the cross-check of inv() below.
template <class P>
__device__ __forceinline__ Fp<P> inv_fermat(const Fp<P>& a) {
  u32 e[8];
  u32 borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) e[i] = sbb(P::m[i], i == 0 ? 2u : 0u, borrow, &borrow);
  return pow_big<P>(a, e);
}

// ---------------------------------------------------------------- inversion
// Optimized binary GCD (T. Pornin, "Optimized Binary GCD for Modular
// Inversion", IACR ePrint 2020/972, Algorithm 2, k = 32).  The classic
// binary GCD (a <- y, b <- m, u <- 1, v <- 0; halve a when even, else
// order a >= b and set a <- (a - b) / 2, u and v following so that
// a = u y and b = v y mod m) is run 31 steps at a time on 64-bit
// approximations of a and b (their low 31 bits and their top 33 bits at
// the length of the larger one); the steps are recorded as a matrix
// (f0 g0; f1 g1) with entries in [-2^31, 2^31], then applied to the full
// values, a <- (a f0 + b g0) / 2^31 and b <- (a f1 + b g1) / 2^31 (exact;
// a negative result is negated with its row), and to u, v with the
// division by 2^31 done modulo m.  ceil((2 * 254 - 1) / 31) = 17 rounds
// bring a to 0 and b to gcd(y, m) = 1, and then v = 1 / y mod m.
namespace inv_detail {
// m < 2^254 for both BN254 fields
constexpr u32 kRounds = (2 * 254 - 1 + 30) / 31;

// x f + y g for x, y < 2^254 and |f|, |g| <= 2^31, as a 9-limb two's
// complement integer (|result| < 2^286)
__device__ __forceinline__ void lin_comb(const u32 x[8], const u32 y[8], long long f, long long g,
                                         u32 r[9]) {
  const bool nf = f < 0, ng = g < 0;
  const u64 mf = (u64)(nf ? -f : f), mg = (u64)(ng ? -g : g);
  u32 p[9], q[9];
  u64 cp = 0, cq = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const u64 tp = (u64)x[i] * mf + cp;
    const u64 tq = (u64)y[i] * mg + cq;
    p[i] = (u32)tp;
    q[i] = (u32)tq;
    cp = tp >> 32;
    cq = tq >> 32;
  }
  p[8] = (u32)cp;
  q[8] = (u32)cq;
  // negate where the factor is negative (two's complement over 9 limbs)
  const u32 sp = 0u - (u32)nf, sq = 0u - (u32)ng;
  u32 bp = (u32)nf, bq = (u32)ng;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const u64 tp = (u64)(p[i] ^ sp) + bp;
    const u64 tq = (u64)(q[i] ^ sq) + bq;
    p[i] = (u32)tp;
    q[i] = (u32)tq;
    bp = (u32)(tp >> 32);
    bq = (u32)(tq >> 32);
  }
  u32 c = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) r[i] = adc(p[i], q[i], c, &c);
}

// x = |r| >> 31 for a 9-limb two's complement r divisible by 2^31; returns
// whether r was negative
__device__ __forceinline__ bool abs_shr31(const u32 r_in[9], u32 x[8]) {
  const bool neg = (r_in[8] >> 31) != 0;
  const u32 s = 0u - (u32)neg;
  u32 r[9], b = (u32)neg;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const u64 t = (u64)(r_in[i] ^ s) + b;
    r[i] = (u32)t;
    b = (u32)(t >> 32);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) x[i] = (r[i] >> 31) | (r[i + 1] << 1);
  return neg;
}

// bits [p, p + 64) of x (p <= 223), the limbs picked by masks (an index
// into x would put it in scratch memory)
__device__ __forceinline__ u64 window64(const u32 x[8], u32 p) {
  const u32 w = p >> 5, sh = p & 31;
  u32 x0 = 0, x1 = 0, x2 = 0;
#pragma unroll
  for (u32 i = 0; i < 8; i++) {
    x0 |= x[i] & (0u - (u32)(i == w));
    x1 |= x[i] & (0u - (u32)(i == w + 1));
    x2 |= x[i] & (0u - (u32)(i == w + 2));
  }
  const u64 lo = ((u64)x1 << 32) | x0;
  return sh ? (lo >> sh) | ((u64)x2 << (64 - sh)) : lo;
}

// bit length of the larger of x and y
__device__ __forceinline__ u32 bit_len2(const u32 x[8], const u32 y[8]) {
  u32 len = 0;
#pragma unroll
  for (u32 i = 0; i < 8; i++) {
    const u32 o = x[i] | y[i];
    if (o) len = 32 * i + 32 - (u32)__builtin_clz(o);
  }
  return len;
}

// (u f + v g) / 2^31 mod m for u, v < m and |f|, |g| <= 2^31: the sum w
// has |w| < m 2^32; m 2^32 is added when it is negative, then
// c = w (-m^-1) mod 2^31 makes w + c m divisible by 2^31, and the quotient
// (< 3m) is brought below m by two conditional subtractions.
template <class P>
__device__ __forceinline__ void mod_update(const u32 u[8], const u32 v[8], long long f, long long g,
                                           u32 out[8]) {
  u32 w[9];
  lin_comb(u, v, f, g, w);
  const u32 neg = 0u - (w[8] >> 31);
  u32 c = 0;
#pragma unroll
  for (int i = 1; i < 9; i++) w[i] = adc(w[i], P::m[i - 1] & neg, c, &c);
  const u32 k = (w[0] * P::np) & 0x7FFFFFFFu;
  u32 s[10];
  u64 carry = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const u64 t = (u64)k * (i < 8 ? P::m[i] : 0u) + w[i] + carry;
    s[i] = (u32)t;
    carry = t >> 32;
  }
  s[9] = (u32)carry;
  u32 t[9];
#pragma unroll
  for (int i = 0; i < 9; i++) t[i] = (s[i] >> 31) | (s[i + 1] << 1);
#pragma unroll
  for (int rep = 0; rep < 2; rep++) {
    u32 d[9], borrow = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) d[i] = sbb(t[i], i < 8 ? P::m[i] : 0u, borrow, &borrow);
    const u32 keep = 0u - borrow;  // t < m: keep t
#pragma unroll
    for (int i = 0; i < 9; i++) t[i] = (t[i] & keep) | (d[i] & ~keep);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = t[i];
}
}  // namespace inv_detail

// 1 / a in Montgomery form: with A = a R the input representative, the GCD
// gives A^-1 = a^-1 R^-1, brought to a^-1 R by two products with R^2.  Same
// value as inv_fermat (the inverse is unique); inv(0) = 0.
template <class P>
__device__ __forceinline__ Fp<P> inv(const Fp<P>& a_in) {
  using namespace inv_detail;
  static_assert((P::m[7] >> 30) == 0 && (P::m[7] >> 29) != 0, "a 254-bit modulus");
  const Fp<P> y = reduce_once<P>(a_in);  // inputs below 2m
  if (y.is_zero()) return y;
  u32 a[8], b[8], u[8], v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    a[i] = y.v[i];
    b[i] = P::m[i];
    u[i] = i == 0 ? 1u : 0u;
    v[i] = 0;
  }
  for (u32 round = 0; round < kRounds; round++) {
    u32 n = bit_len2(a, b);
    n = n < 64 ? 64 : n;
    u64 A = (u64)(a[0] & 0x7FFFFFFFu) | (window64(a, n - 33) << 31);
    u64 B = (u64)(b[0] & 0x7FFFFFFFu) | (window64(b, n - 33) << 31);
    // the matrix entries in 32-bit two's complement: |f| + |g| <= 2^31 in
    // each row, so the one value that does not fit, +2^31, comes with a 0
    // beside it; read back as -2^31 it negates its row's result, which the
    // sign correction below undoes (a zero result stays zero)
    u32 f0 = 1, g0 = 0, f1 = 0, g1 = 1;
#pragma unroll 1
    for (int j = 0; j < 31; j++) {
      if (A & 1) {
        if (A < B) {
          const u64 tA = A;
          A = B;
          B = tA;
          const u32 tf = f0, tg = g0;
          f0 = f1;
          g0 = g1;
          f1 = tf;
          g1 = tg;
        }
        A -= B;
        f0 -= f1;
        g0 -= g1;
      }
      A >>= 1;
      f1 <<= 1;
      g1 <<= 1;
    }
    long long F0 = (int)f0, G0 = (int)g0, F1 = (int)f1, G1 = (int)g1;
    u32 ra[9], rb[9], na[8], nb[8];
    lin_comb(a, b, F0, G0, ra);
    lin_comb(a, b, F1, G1, rb);
    if (abs_shr31(ra, na)) {
      F0 = -F0;
      G0 = -G0;
    }
    if (abs_shr31(rb, nb)) {
      F1 = -F1;
      G1 = -G1;
    }
    u32 nu[8], nv[8];
    mod_update<P>(u, v, F0, G0, nu);
    mod_update<P>(u, v, F1, G1, nv);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      a[i] = na[i];
      b[i] = nb[i];
      u[i] = nu[i];
      v[i] = nv[i];
    }
  }
  Fp<P> r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.v[i] = v[i];
  return mul<P>(mul<P>(r, Fp<P>::r2()), Fp<P>::r2());
}

// ------------------------------------------------------------------ G1 curve
struct G1Aff {
  Fq x, y;
  __device__ __forceinline__ bool is_identity() const { return y.is_zero(); }
};
struct G1Xyzz {
  Fq x, y, zz, zzz;
  __device__ __forceinline__ static G1Xyzz identity() {
    G1Xyzz r;
    r.x = Fq::zero();
    r.y = Fq::zero();
    r.zz = Fq::zero();
    r.zzz = Fq::one();
    return r;
  }
  __device__ __forceinline__ bool is_identity() const { return zz.is_zero(); }
};

__device__ __forceinline__ G1Aff load_aff(const u32* p) {
  G1Aff r;
  r.x = load_fe<Fq>(p);
  r.y = load_fe<Fq>(p + 8);
  return r;
}
__device__ __forceinline__ void store_aff(u32* p, const G1Aff& a) {
  store_fe<Fq>(p, a.x);
  store_fe<Fq>(p + 8, a.y);
}
__device__ __forceinline__ G1Xyzz load_xyzz(const u32* p) {
  G1Xyzz r;
  r.x = load_fe<Fq>(p);
  r.y = load_fe<Fq>(p + 8);
  r.zz = load_fe<Fq>(p + 16);
  r.zzz = load_fe<Fq>(p + 24);
  return r;
}
__device__ __forceinline__ void store_xyzz(u32* p, const G1Xyzz& a) {
  store_fe<Fq>(p, a.x);
  store_fe<Fq>(p + 8, a.y);
  store_fe<Fq>(p + 16, a.zz);
  store_fe<Fq>(p + 24, a.zzz);
}

__device__ __forceinline__ G1Xyzz aff_to_xyzz(const G1Aff& a) {
  if (a.is_identity()) return G1Xyzz::identity();
  G1Xyzz r;
  r.x = a.x;
  r.y = a.y;
  r.zz = Fq::one();
  r.zzz = Fq::one();
  return r;
}

// mdbl-2008-s-1: 2 * affine point
__device__ __forceinline__ G1Xyzz aff_dbl(const G1Aff& p) {
  typedef FqParams Q;
  Fq u = dbl<Q>(p.y);
  Fq v = sqr<Q>(u);
  Fq w = mul<Q>(u, v);
  Fq s = mul<Q>(p.x, v);
  Fq x2 = sqr<Q>(p.x);
  Fq m = add<Q>(dbl<Q>(x2), x2);
  Fq x3 = sub<Q>(sub<Q>(sqr<Q>(m), s), s);
  Fq y3 = sub<Q>(mul<Q>(m, sub<Q>(s, x3)), mul<Q>(w, p.y));
  G1Xyzz r;
  r.x = x3;
  r.y = y3;
  r.zz = v;
  r.zzz = w;
  return r;
}

// dbl-2008-s-1
__device__ __forceinline__ G1Xyzz xyzz_dbl(const G1Xyzz& p) {
  typedef FqParams Q;
  if (p.is_identity()) return p;
  Fq u = dbl<Q>(p.y);
  Fq v = sqr<Q>(u);
  Fq w = mul<Q>(u, v);
  Fq s = mul<Q>(p.x, v);
  Fq x2 = sqr<Q>(p.x);
  Fq m = add<Q>(dbl<Q>(x2), x2);
  Fq x3 = sub<Q>(sub<Q>(sqr<Q>(m), s), s);
  Fq y3 = sub<Q>(mul<Q>(m, sub<Q>(s, x3)), mul<Q>(w, p.y));
  G1Xyzz r;
  r.x = x3;
  r.y = y3;
  r.zz = mul<Q>(v, p.zz);
  r.zzz = mul<Q>(w, p.zzz);
  return r;
}

// madd-2008-s: XYZZ + affine
__device__ __forceinline__ G1Xyzz xyzz_madd(const G1Xyzz& p, const G1Aff& q) {
  typedef FqParams Q;
  if (q.is_identity()) return p;
  if (p.is_identity()) return aff_to_xyzz(q);
  Fq u2 = mul<Q>(q.x, p.zz);
  Fq s2 = mul<Q>(q.y, p.zzz);
  Fq P_ = sub<Q>(u2, p.x);
  Fq R_ = sub<Q>(s2, p.y);
  if (P_.is_zero() && R_.is_zero()) return aff_dbl(q);
  Fq pp = sqr<Q>(P_);
  Fq ppp = mul<Q>(P_, pp);
  Fq qq = mul<Q>(p.x, pp);
  Fq x3 = sub<Q>(sub<Q>(sub<Q>(sqr<Q>(R_), ppp), qq), qq);
  Fq y3 = sub<Q>(mul<Q>(R_, sub<Q>(qq, x3)), mul<Q>(p.y, ppp));
  G1Xyzz r;
  r.x = x3;
  r.y = y3;
  r.zz = mul<Q>(p.zz, pp);
  r.zzz = mul<Q>(p.zzz, ppp);
  return r;
}

// add-2008-s: XYZZ + XYZZ
__device__ __forceinline__ G1Xyzz xyzz_add(const G1Xyzz& p, const G1Xyzz& q) {
  typedef FqParams Q;
  if (p.is_identity()) return q;
  if (q.is_identity()) return p;
  Fq u1 = mul<Q>(p.x, q.zz);
  Fq u2 = mul<Q>(q.x, p.zz);
  Fq s1 = mul<Q>(p.y, q.zzz);
  Fq s2 = mul<Q>(q.y, p.zzz);
  Fq P_ = sub<Q>(u2, u1);
  Fq R_ = sub<Q>(s2, s1);
  if (P_.is_zero() && R_.is_zero()) return xyzz_dbl(p);
  Fq pp = sqr<Q>(P_);
  Fq ppp = mul<Q>(P_, pp);
  Fq qq = mul<Q>(u1, pp);
  Fq x3 = sub<Q>(sub<Q>(sub<Q>(sqr<Q>(R_), ppp), qq), qq);
  Fq y3 = sub<Q>(mul<Q>(R_, sub<Q>(qq, x3)), mul<Q>(s1, ppp));
  G1Xyzz r;
  r.x = x3;
  r.y = y3;
  r.zz = mul<Q>(mul<Q>(p.zz, q.zz), pp);
  r.zzz = mul<Q>(mul<Q>(p.zzz, q.zzz), ppp);
  return r;
}

// ---------------------------------------------- group-of-four broadcast
// v as held by lane 4 * (lane / 4) + K, for every lane of each group of four
// consecutive lanes: a DPP quad permutation (one VALU move, no LDS).  Every
// lane of a group that uses it must be active.
template <int K>
__device__ __forceinline__ u32 quad_bcast(u32 v) {
  static_assert(K >= 0 && K < 4, "lane in the group");
#if defined(IMP_EMU)
  return (u32)::emu::quad_shfl(v, K);
#elif defined(__HIP_DEVICE_COMPILE__)
  return (u32)__builtin_amdgcn_mov_dpp((int)v, K * 0x55, 0xF, 0xF, false);
#else
  return v;
#endif
}

__device__ __forceinline__ G1Aff aff_neg(const G1Aff& a) {
  G1Aff r;
  r.x = a.x;
  r.y = neg<FqParams>(a.y);
  return r;
}

// z-scaling to affine; identity -> (0, 0).  all_lanes: every lane of the wave
// converts (Fermat's fixed exponent chain avoids the divergence of the binary
// inversion, which is faster for a single lane)
__device__ __forceinline__ G1Aff xyzz_to_aff(const G1Xyzz& p, bool all_lanes = false) {
  typedef FqParams Q;
  G1Aff r;
  if (p.is_identity()) {
    r.x = Fq::zero();
    r.y = Fq::zero();
    return r;
  }
  Fq a = all_lanes ? inv_fermat<Q>(p.zzz) : inv<Q>(p.zzz);
  Fq b = sqr<Q>(mul<Q>(p.zz, a));
  r.x = mul<Q>(p.x, b);
  r.y = mul<Q>(p.y, a);
  return r;
}

// k * P, k a 256-bit plain integer (double-and-add from the top bit)
__device__ __forceinline__ G1Xyzz scalar_mul(const G1Aff& p, const u32 k[8]) {
  G1Xyzz acc = G1Xyzz::identity();
  for (int w = 7; w >= 0; w--) {
    for (int i = 31; i >= 0; i--) {
      acc = xyzz_dbl(acc);
      if ((k[w] >> i) & 1) acc = xyzz_madd(acc, p);
    }
  }
  return acc;
}

// k * P for a small k (< 2^32), XYZZ input
__device__ __forceinline__ G1Xyzz xyzz_mul_u32(const G1Xyzz& p, u32 k) {
  G1Xyzz acc = G1Xyzz::identity();
  for (int i = 31; i >= 0; i--) {
    acc = xyzz_dbl(acc);
    if ((k >> i) & 1) acc = xyzz_add(acc, p);
  }
  return acc;
}

// ------------------------------------------------------------- seeded inputs
// Counter-based SplitMix64 generator.  The same definition lives in the CPU
// checker so both sides see identical bytes for a (seed, index).
__host__ __device__ __forceinline__ u64 smix(u64 z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform Fr (its Montgomery representation is uniform in [0, r))
__device__ __forceinline__ Fr gen_uniform_fr(u64 seed, u64 i) {
  u64 key = smix(seed) ^ (i * 0xD1B54A32D192ED03ull);
  for (u64 a = 0;; a++) {
    u64 w[4];
#pragma unroll
    for (int j = 0; j < 4; j++) w[j] = smix(key + 4 * a + (u64)j);
    w[3] &= 0x3FFFFFFFFFFFFFFFull;
    Fr r;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      r.v[2 * j] = (u32)w[j];
      r.v[2 * j + 1] = (u32)(w[j] >> 32);
    }
    // accept if r < m
    u32 borrow = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) (void)sbb(r.v[j], FrParams::m[j], borrow, &borrow);
    if (borrow) return r;
  }
}

}  // namespace imp
