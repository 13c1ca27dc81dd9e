// host_field.h -- BN254 Fr Montgomery arithmetic on the host CPU, for setup
// tables that a caller asks for in host memory (gen_pq_omegas): 4 x 64-bit
// limbs, CIOS product (Koc, Acar, Kaliski 1996) with unsigned __int128
// partial products, canonical results.  Same values as the device code (the
// Montgomery form with R = 2^256 of a field element is unique).
#pragma once
#include <stdint.h>
#include <string.h>

#include "field.h"

namespace imp {
namespace host {

struct Fe {
  uint64_t v[4];
};

inline Fe load(const uint32_t* p) {
  Fe a;
  for (int i = 0; i < 4; i++) a.v[i] = (uint64_t)p[2 * i] | ((uint64_t)p[2 * i + 1] << 32);
  return a;
}

inline void store(uint32_t* p, const Fe& a) {
  for (int i = 0; i < 4; i++) {
    p[2 * i] = (uint32_t)a.v[i];
    p[2 * i + 1] = (uint32_t)(a.v[i] >> 32);
  }
}

template <class P>
inline Fe modulus() {
  Fe m;
  for (int i = 0; i < 4; i++) m.v[i] = (uint64_t)P::m[2 * i] | ((uint64_t)P::m[2 * i + 1] << 32);
  return m;
}

template <class P>
inline Fe one() {
  Fe r;
  for (int i = 0; i < 4; i++) r.v[i] = (uint64_t)P::r1[2 * i] | ((uint64_t)P::r1[2 * i + 1] << 32);
  return r;
}

// -m^-1 mod 2^64 by Newton iteration (each step doubles the correct bits)
template <class P>
inline uint64_t neg_inv64() {
  const uint64_t m0 = modulus<P>().v[0];
  uint64_t x = 1;
  for (int i = 0; i < 7; i++) x *= 2 - m0 * x;
  return (uint64_t)0 - x;
}

// a * b * 2^-256 mod m, canonical
template <class P>
inline Fe mul(const Fe& a, const Fe& b) {
  typedef unsigned __int128 u128;
  const Fe m = modulus<P>();
  const uint64_t np = neg_inv64<P>();
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    u128 c = 0;
    for (int j = 0; j < 4; j++) {
      c = (u128)t[j] + (u128)a.v[j] * b.v[i] + (uint64_t)(c >> 64);
      t[j] = (uint64_t)c;
    }
    u128 s = (u128)t[4] + (uint64_t)(c >> 64);
    t[4] = (uint64_t)s;
    t[5] = (uint64_t)(s >> 64);
    const uint64_t q = t[0] * np;
    c = (u128)t[0] + (u128)q * m.v[0];
    for (int j = 1; j < 4; j++) {
      c = (u128)t[j] + (u128)q * m.v[j] + (uint64_t)(c >> 64);
      t[j - 1] = (uint64_t)c;
    }
    s = (u128)t[4] + (uint64_t)(c >> 64);
    t[3] = (uint64_t)s;
    t[4] = t[5] + (uint64_t)(s >> 64);
  }
  // t < 2m: one conditional subtraction
  Fe r, d;
  uint64_t borrow = 0;
  for (int j = 0; j < 4; j++) {
    const u128 x = (u128)t[j] - m.v[j] - borrow;
    d.v[j] = (uint64_t)x;
    borrow = (uint64_t)(x >> 64) & 1;
  }
  const bool ge = t[4] != 0 || borrow == 0;
  for (int j = 0; j < 4; j++) r.v[j] = ge ? d.v[j] : t[j];
  return r;
}

// a^e for a 64-bit exponent (left to right square and multiply)
template <class P>
inline Fe pow(const Fe& a, uint64_t e) {
  Fe r = one<P>();
  for (int k = 63; k >= 0; k--) {
    r = mul<P>(r, r);
    if ((e >> k) & 1) r = mul<P>(r, a);
  }
  return r;
}

}  // namespace host
}  // namespace imp
