/*
 * oracle.h -- CPU restatement of Impeller's BN254 NTT / MSM / fused-gate path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (impeller_amd/, include/)
 * links, loads or calls this library; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker.
 *
 * Data layout is the reference's: a field element is 8 little-endian u32 limbs
 * (== 4 little-endian u64 limbs, halo2curves' memory layout), Montgomery form
 * with R = 2^256, canonical (< modulus).  A G1 affine point is x || y (64 B),
 * identity = (0, 0).  An "integer" argument is a plain 256-bit little-endian
 * number (not Montgomery).
 *
 * Field id: ORC_FR = 0 (BN254 scalar field), ORC_FQ = 1 (BN254 base field).
 */
#ifndef IMPELLER_ORACLE_H
#define IMPELLER_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_FR = 0, ORC_FQ = 1 };

/* ---- field (core/src/common/mont/src/field.cuh:29-158 host semantics) ---- */
void orc_mul(int f, uint32_t *r, const uint32_t *a, const uint32_t *b);
void orc_sqr(int f, uint32_t *r, const uint32_t *a);
void orc_add(int f, uint32_t *r, const uint32_t *a, const uint32_t *b);
void orc_sub(int f, uint32_t *r, const uint32_t *a, const uint32_t *b);
void orc_neg(int f, uint32_t *r, const uint32_t *a);
void orc_pow(int f, uint32_t *r, const uint32_t *a, const uint32_t *e_int);
void orc_inv(int f, uint32_t *r, const uint32_t *a);
void orc_from_number(int f, uint32_t *r, const uint32_t *n_int);
void orc_to_number(int f, uint32_t *r, const uint32_t *a);
void orc_one(int f, uint32_t *r);
void orc_modulus(int f, uint32_t *r);
/* elementwise bulk ops over n elements */
void orc_vec_mul(int f, uint32_t *r, const uint32_t *a, const uint32_t *b, uint64_t n);
void orc_vec_scale(int f, uint32_t *x, uint64_t n, const uint32_t *s);

/* ---- seeded inputs (shared definition with the product's device generator) */
/* kind 0: uniform in [0, r) (the value is the Montgomery representation);
 * kind 1: selector-like, 90% zero, else one;
 * kind 2: small integer value < 2^16 (stored in Montgomery form). */
void orc_gen_scalars(int kind, uint64_t seed, uint64_t start, uint64_t n, uint32_t *out);

/* ---- NTT (halo2 best_fft restated: core/tests/msm_halo2.rs:221-324) ---- */
void orc_best_fft(uint32_t *a, const uint32_t *omega, uint32_t log_n, int threads);
/* Horner evaluation sum_i coeffs[i] * point^i (pins single outputs of a
 * full-size NTT: y_k = p(omega^k) for a true root of unity) */
void orc_poly_eval(const uint32_t *coeffs, uint64_t n, const uint32_t *point, uint32_t *r);
/* twiddle table omega^i, i < len (ssip_precompute semantics, precompute.cuh:36-84) */
void orc_powers(uint32_t *out, uint64_t len, const uint32_t *base);
/* x[i] *= powers[slice.to_origin(i) % power_num]  (distribute_powers.cuh:5-10),
 * x addressed as a PolyPtr view (iter.cuh:309-317) */
void orc_distribute_powers(uint32_t *x, uint64_t len, uint64_t rotate, uint64_t offset,
                           uint64_t whole_len, const uint32_t *powers, uint64_t power_num,
                           uint64_t s_begin, uint64_t s_slice_len, uint64_t s_len);

/* ---- G1 (core/src/common/curve/src/curve_xyzz.cuh:169-418) ---- */
/* XYZZ points are x || y || zz || zzz (128 B) */
void orc_xyzz_madd(uint32_t *r, const uint32_t *p_xyzz, const uint32_t *q_aff);
void orc_xyzz_add(uint32_t *r, const uint32_t *p_xyzz, const uint32_t *q_xyzz);
void orc_xyzz_dbl(uint32_t *r, const uint32_t *p_xyzz);
void orc_xyzz_to_affine(uint32_t *r_aff, const uint32_t *p_xyzz);
void orc_affine_to_xyzz(uint32_t *r, const uint32_t *p_aff);
int orc_affine_on_curve(const uint32_t *p_aff);
int orc_xyzz_on_curve(const uint32_t *p_xyzz);
void orc_g1_generator(uint32_t *r_aff);
/* k * P (k a 256-bit integer), affine in, affine out */
void orc_g1_mul(uint32_t *r_aff, const uint32_t *p_aff, const uint32_t *k_int);
/* bases[i] = to_number(k_i) * G where k_i = orc_gen_scalars(0, seed, i) */
void orc_gen_bases(uint64_t seed, uint64_t start, uint64_t n, uint32_t *bases_aff, int threads);

/* ---- MSM ---- */
/* halo2 best_multiexp restated (msm_halo2.rs:63-209), result normalised to affine */
void orc_best_multiexp(const uint32_t *scalars, const uint32_t *bases_aff, uint64_t n,
                       int threads, uint32_t *r_aff);
/* closed form for bases k_i*G: (sum_i s_i k_i mod r) * G */
void orc_msm_closed_form(const uint32_t *scalars, const uint32_t *k, uint64_t n, uint32_t *r_aff);

/* ---- fused gate over Fr: out = a*b*c - d + q*e*(1 - q) ---- */
void orc_gate(uint32_t *out, const uint32_t *a, const uint32_t *b, const uint32_t *c,
              const uint32_t *d, const uint32_t *q, const uint32_t *e, uint64_t n);

/* ---- polynomial library (core/src/poly/src/poly.h:6-30) ---- */
/* op 0 add, 1 sub, 2 mul over mixed lengths (poly_basic.cu:4-54); r holds max(na, nb) */
void orc_poly_binary(int op, uint32_t *r, const uint32_t *a, uint64_t na, const uint32_t *b,
                     uint64_t nb);
/* halo2 kate_division, padded with q[n-1] = 0 */
void orc_kate_division(const uint32_t *p, uint64_t n, const uint32_t *b, uint32_t *q);
/* ff::BatchInvert in place (zeros stay zero) */
void orc_batch_invert(uint32_t *x, uint64_t n);
/* inclusive prefix products in place */
void orc_scan_mul(uint32_t *x, uint64_t n);
/* lookup permutation (permute.cuh); returns the number of input values missing from the table */
uint64_t orc_permute(const uint32_t *input, const uint32_t *table, uint64_t usable,
                     uint32_t *res_input, uint32_t *res_table);

#ifdef __cplusplus
}
#endif
#endif
