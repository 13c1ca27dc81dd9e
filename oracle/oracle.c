/*
 * oracle.c -- CPU restatement of Impeller's BN254 NTT / MSM / fused-gate path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never shipped or
 * measured as the product.  Arithmetic is on 4 x u64 limbs with 128-bit
 * products; the Montgomery representation (R = 2^256, canonical < modulus) is
 * byte-identical to the reference's 8 x u32 limbs.
 *
 * What each part restates:
 *   field  : core/src/common/mont/src/field.cuh:29-158 (host CIOS, canonical
 *            output), :970-1175 (mul/add/sub/neg/pow/invert = pow(m-2)),
 *            constants core/src/common/mont/src/bn254_fr.cuh:18-89,
 *            bn254_fq.cuh:10-84
 *   curve  : core/src/common/curve/src/curve_xyzz.cuh:169-418 (EFD XYZZ:
 *            madd-2008-s, add-2008-s, dbl-2008-s-1, mdbl-2008-s-1, z-scaling),
 *            b = 3 (bn254.cuh:11-18)
 *   NTT    : halo2 best_fft as vendored at core/tests/msm_halo2.rs:221-324
 *            (bit reversal, n/2 twiddle table, recursive DIT butterflies)
 *   powers : core/src/ntt/src/distribute_powers.cuh:5-10 and the PolyPtr view
 *            of core/src/common/iter/src/iter.cuh:309-317
 *   MSM    : halo2 multiexp_serial / best_multiexp as vendored at
 *            core/tests/msm_halo2.rs:63-209
 *   gate   : element-wise semantics of the generated fused kernels
 *            (core/src/fused_kernels/kernel_gen.rs:460-555)
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef struct { uint64_t v[4]; } fe;

typedef struct {
    fe m;        /* modulus */
    uint64_t np; /* -m^-1 mod 2^64 */
    fe r1;       /* R mod m */
    fe r2;       /* R^2 mod m */
} fparams;

/* constants: bn254_fr.cuh:21-25, bn254_fq.cuh:13-17 (little-endian u64) */
static fparams FP[2] = {
    {{{0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull, 0x30644e72e131a029ull}},
     0,
     {{0xac96341c4ffffffbull, 0x36fc76959f60cd29ull, 0x666ea36f7879462eull, 0x0e0a77c19a07df2full}},
     {{0x1bb8e645ae216da7ull, 0x53fe3ab1e35c59e3ull, 0x8c49833d53bb8085ull, 0x0216d0b17f4e44a5ull}}},
    {{{0x3c208c16d87cfd47ull, 0x97816a916871ca8dull, 0xb85045b68181585dull, 0x30644e72e131a029ull}},
     0,
     {{0xd35d438dc58f0d9dull, 0x0a78eb28f5c70b3dull, 0x666ea36f7879462cull, 0x0e0a77c19a07df2full}},
     {{0xf32cfc5b538afa89ull, 0xb5e71911d44501fbull, 0x47ab1eff0a417ff6ull, 0x06d89f71cab8351full}}},
};

static void params_init(void) {
    static int done = 0;
    if (done) return;
    for (int f = 0; f < 2; f++) {
        /* Newton iteration for m^-1 mod 2^64 (m odd) */
        uint64_t m0 = FP[f].m.v[0], inv = 1;
        for (int i = 0; i < 7; i++) inv *= 2 - m0 * inv;
        FP[f].np = (uint64_t)0 - inv;
    }
    done = 1;
}

static inline fe load(const uint32_t *p) {
    fe r;
    memcpy(r.v, p, 32);
    return r;
}
static inline void store(uint32_t *p, const fe *a) { memcpy(p, a->v, 32); }

static inline int is_zero(const fe *a) { return (a->v[0] | a->v[1] | a->v[2] | a->v[3]) == 0; }
static inline int fe_eq(const fe *a, const fe *b) {
    return a->v[0] == b->v[0] && a->v[1] == b->v[1] && a->v[2] == b->v[2] && a->v[3] == b->v[3];
}
/* a >= b as integers */
static inline int geq(const fe *a, const fe *b) {
    for (int i = 3; i >= 0; i--) {
        if (a->v[i] != b->v[i]) return a->v[i] > b->v[i];
    }
    return 1;
}
static inline uint64_t sub_raw(fe *r, const fe *a, const fe *b) {
    uint64_t borrow = 0;
    for (int i = 0; i < 4; i++) {
        u128 d = (u128)a->v[i] - b->v[i] - borrow;
        r->v[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) & 1;
    }
    return borrow;
}
static inline uint64_t add_raw(fe *r, const fe *a, const fe *b) {
    u128 c = 0;
    for (int i = 0; i < 4; i++) {
        c += (u128)a->v[i] + b->v[i];
        r->v[i] = (uint64_t)c;
        c >>= 64;
    }
    return (uint64_t)c;
}

/* CIOS Montgomery product a*b*R^-1 mod m, canonical output (field.cuh:91-134) */
static inline void fmul(const fparams *P, fe *r, const fe *a, const fe *b) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        u128 c = 0;
        for (int j = 0; j < 4; j++) {
            c += (u128)a->v[j] * b->v[i] + t[j];
            t[j] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[4] = (uint64_t)c;
        t[5] = (uint64_t)(c >> 64);
        uint64_t k = t[0] * P->np;
        c = (u128)k * P->m.v[0] + t[0];
        c >>= 64;
        for (int j = 1; j < 4; j++) {
            c += (u128)k * P->m.v[j] + t[j];
            t[j - 1] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[3] = (uint64_t)c;
        t[4] = t[5] + (uint64_t)(c >> 64);
    }
    fe res = {{t[0], t[1], t[2], t[3]}};
    if (t[4] || geq(&res, &P->m)) sub_raw(&res, &res, &P->m);
    *r = res;
}
static inline void fadd(const fparams *P, fe *r, const fe *a, const fe *b) {
    fe s;
    uint64_t c = add_raw(&s, a, b);
    if (c || geq(&s, &P->m)) sub_raw(&s, &s, &P->m);
    *r = s;
}
static inline void fsub(const fparams *P, fe *r, const fe *a, const fe *b) {
    fe s;
    if (sub_raw(&s, a, b)) add_raw(&s, &s, &P->m);
    *r = s;
}
static inline void fneg(const fparams *P, fe *r, const fe *a) {
    if (is_zero(a)) {
        *r = *a;
        return;
    }
    sub_raw(r, &P->m, a);
}
/* square-and-multiply from the top bit (field.cuh:1133-1169) */
static void fpow(const fparams *P, fe *r, const fe *a, const fe *e) {
    fe acc = P->r1;
    int started = 0;
    for (int i = 255; i >= 0; i--) {
        if (started) fmul(P, &acc, &acc, &acc);
        if ((e->v[i >> 6] >> (i & 63)) & 1) {
            fmul(P, &acc, &acc, a);
            started = 1;
        }
    }
    *r = acc;
}
static void finv(const fparams *P, fe *r, const fe *a) {
    fe e = P->m, two = {{2, 0, 0, 0}};
    sub_raw(&e, &e, &two);
    fpow(P, r, a, &e);
}

/* ------------------------------------------------------------------ API */
void orc_mul(int f, uint32_t *r, const uint32_t *a, const uint32_t *b) {
    params_init();
    fe x = load(a), y = load(b), z;
    fmul(&FP[f], &z, &x, &y);
    store(r, &z);
}
void orc_sqr(int f, uint32_t *r, const uint32_t *a) { orc_mul(f, r, a, a); }
void orc_add(int f, uint32_t *r, const uint32_t *a, const uint32_t *b) {
    fe x = load(a), y = load(b), z;
    fadd(&FP[f], &z, &x, &y);
    store(r, &z);
}
void orc_sub(int f, uint32_t *r, const uint32_t *a, const uint32_t *b) {
    fe x = load(a), y = load(b), z;
    fsub(&FP[f], &z, &x, &y);
    store(r, &z);
}
void orc_neg(int f, uint32_t *r, const uint32_t *a) {
    fe x = load(a), z;
    fneg(&FP[f], &z, &x);
    store(r, &z);
}
void orc_pow(int f, uint32_t *r, const uint32_t *a, const uint32_t *e_int) {
    params_init();
    fe x = load(a), e = load(e_int), z;
    fpow(&FP[f], &z, &x, &e);
    store(r, &z);
}
void orc_inv(int f, uint32_t *r, const uint32_t *a) {
    params_init();
    fe x = load(a), z;
    finv(&FP[f], &z, &x);
    store(r, &z);
}
void orc_from_number(int f, uint32_t *r, const uint32_t *n_int) {
    params_init();
    fe x = load(n_int), z;
    fmul(&FP[f], &z, &x, &FP[f].r2);
    store(r, &z);
}
void orc_to_number(int f, uint32_t *r, const uint32_t *a) {
    params_init();
    fe x = load(a), one = {{1, 0, 0, 0}}, z;
    fmul(&FP[f], &z, &x, &one);
    store(r, &z);
}
void orc_one(int f, uint32_t *r) { store(r, &FP[f].r1); }
void orc_modulus(int f, uint32_t *r) { store(r, &FP[f].m); }

void orc_vec_mul(int f, uint32_t *r, const uint32_t *a, const uint32_t *b, uint64_t n) {
    params_init();
    for (uint64_t i = 0; i < n; i++) {
        fe x = load(a + 8 * i), y = load(b + 8 * i), z;
        fmul(&FP[f], &z, &x, &y);
        store(r + 8 * i, &z);
    }
}
void orc_vec_scale(int f, uint32_t *x, uint64_t n, const uint32_t *s) {
    params_init();
    fe k = load(s);
    for (uint64_t i = 0; i < n; i++) {
        fe y = load(x + 8 * i);
        fmul(&FP[f], &y, &y, &k);
        store(x + 8 * i, &y);
    }
}

/* ------------------------------------------------------- seeded inputs */
static inline uint64_t smix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
/* uniform in [0, r): rejection sampling of 254-bit words */
static fe gen_uniform(uint64_t seed, uint64_t i) {
    const fe *m = &FP[0].m;
    uint64_t key = smix(seed) ^ (i * 0xD1B54A32D192ED03ull);
    for (uint64_t a = 0;; a++) {
        fe w;
        for (int j = 0; j < 4; j++) w.v[j] = smix(key + 4 * a + (uint64_t)j);
        w.v[3] &= 0x3FFFFFFFFFFFFFFFull;
        if (!geq(&w, m)) return w;
    }
}
void orc_gen_scalars(int kind, uint64_t seed, uint64_t start, uint64_t n, uint32_t *out) {
    params_init();
    const fparams *P = &FP[0];
    for (uint64_t t = 0; t < n; t++) {
        uint64_t i = start + t;
        fe v;
        if (kind == 0) {
            v = gen_uniform(seed, i);
        } else {
            uint64_t h = smix(smix(seed) ^ (i * 0xD1B54A32D192ED03ull) ^ 0x5DEECE66Dull);
            if (kind == 1) {
                /* 90% zero, 10% one */
                if ((h % 10) == 0) v = P->r1;
                else memset(&v, 0, sizeof v);
            } else {
                fe sm = {{h & 0xFFFFull, 0, 0, 0}};
                fmul(P, &v, &sm, &P->r2);
            }
        }
        store(out + 8 * t, &v);
    }
}

/* ------------------------------------------------------------------ NTT */
static uint64_t bitrev(uint64_t x, uint32_t l) {
    uint64_t r = 0;
    for (uint32_t i = 0; i < l; i++) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

void orc_powers(uint32_t *out, uint64_t len, const uint32_t *base) {
    params_init();
    const fparams *P = &FP[0];
    fe w = P->r1, b = load(base);
    for (uint64_t i = 0; i < len; i++) {
        store(out + 8 * i, &w);
        fmul(P, &w, &w, &b);
    }
}

void orc_poly_eval(const uint32_t *coeffs, uint64_t n, const uint32_t *point, uint32_t *r) {
    params_init();
    const fparams *P = &FP[0];
    fe acc = {{0, 0, 0, 0}}, x = load(point);
    for (uint64_t i = n; i-- > 0;) {
        fe c = load(coeffs + 8 * i);
        fmul(P, &acc, &acc, &x);
        fadd(P, &acc, &acc, &c);
    }
    store(r, &acc);
}

typedef struct {
    fe *a;
    uint64_t n, chunk;
    const fe *tw;
    int depth_par;
} fft_job;

/* DIT butterflies of one sub-transform after its halves are done:
 * msm_halo2.rs:297-323 (twiddle 1 at i = 0, then twiddles[(i+1)*chunk]) */
static void fft_merge(fe *a, uint64_t n, uint64_t chunk, const fe *tw) {
    const fparams *P = &FP[0];
    uint64_t h = n / 2;
    fe t = a[h];
    a[h] = a[0];
    fadd(P, &a[0], &a[0], &t);
    fsub(P, &a[h], &a[h], &t);
    for (uint64_t i = 1; i < h; i++) {
        fmul(P, &t, &a[h + i], &tw[i * chunk]);
        a[h + i] = a[i];
        fadd(P, &a[i], &a[i], &t);
        fsub(P, &a[h + i], &a[h + i], &t);
    }
}

static void *fft_rec(void *arg);
static void fft_recursive(fe *a, uint64_t n, uint64_t chunk, const fe *tw, int depth_par) {
    const fparams *P = &FP[0];
    if (n == 2) {
        fe t = a[1];
        a[1] = a[0];
        fadd(P, &a[0], &a[0], &t);
        fsub(P, &a[1], &a[1], &t);
        return;
    }
    if (depth_par > 0) {
        /* multicore::join of the two halves (msm_halo2.rs:292-295) */
        fft_job right = {a + n / 2, n / 2, chunk * 2, tw, depth_par - 1};
        pthread_t th;
        if (pthread_create(&th, NULL, fft_rec, &right) == 0) {
            fft_recursive(a, n / 2, chunk * 2, tw, depth_par - 1);
            pthread_join(th, NULL);
        } else {
            fft_recursive(a, n / 2, chunk * 2, tw, 0);
            fft_recursive(a + n / 2, n / 2, chunk * 2, tw, 0);
        }
    } else {
        fft_recursive(a, n / 2, chunk * 2, tw, 0);
        fft_recursive(a + n / 2, n / 2, chunk * 2, tw, 0);
    }
    fft_merge(a, n, chunk, tw);
}
static void *fft_rec(void *arg) {
    fft_job *j = (fft_job *)arg;
    fft_recursive(j->a, j->n, j->chunk, j->tw, j->depth_par);
    return NULL;
}

static uint32_t log2_floor_u(uint64_t x) {
    uint32_t p = 0;
    while ((2ull << p) <= x) p++;
    return p;
}

void orc_best_fft(uint32_t *a_raw, const uint32_t *omega_raw, uint32_t log_n, int threads) {
    params_init();
    const fparams *P = &FP[0];
    uint64_t n = 1ull << log_n;
    fe *a = (fe *)a_raw; /* same bytes: 4 LE u64 == 8 LE u32 */
    fe omega = load(omega_raw);
    if (threads < 1) threads = 1;
    uint32_t log_threads = log2_floor_u((uint64_t)threads);

    for (uint64_t k = 0; k < n; k++) {
        uint64_t rk = bitrev(k, log_n);
        if (k < rk) {
            fe t = a[k];
            a[k] = a[rk];
            a[rk] = t;
        }
    }
    if (log_n == 0) return;
    uint64_t h = n / 2;
    fe *tw = (fe *)malloc(sizeof(fe) * (h ? h : 1));
    fe w = P->r1;
    for (uint64_t i = 0; i < h; i++) {
        tw[i] = w;
        fmul(P, &w, &w, &omega);
    }
    if (log_n <= log_threads) {
        /* iterative variant (msm_halo2.rs:254-283): same butterflies */
        uint64_t chunk = 2, twc = n / 2;
        for (uint32_t s = 0; s < log_n; s++) {
            for (uint64_t base = 0; base < n; base += chunk) fft_merge(a + base, chunk, twc, tw);
            chunk *= 2;
            twc /= 2;
        }
    } else {
        fft_recursive(a, n, 1, tw, (int)log_threads);
    }
    free(tw);
}

/* PolyPtr logical index -> physical index (iter.cuh:309-317, SliceIterator::operator[]) */
static uint64_t view_index(uint64_t i, uint64_t len, uint64_t rotate, uint64_t offset,
                           uint64_t whole_len) {
    if (rotate != 0) offset = len - rotate;
    return (offset + i) % whole_len;
}

void orc_distribute_powers(uint32_t *x, uint64_t len, uint64_t rotate, uint64_t offset,
                           uint64_t whole_len, const uint32_t *powers, uint64_t power_num,
                           uint64_t s_begin, uint64_t s_slice_len, uint64_t s_len) {
    params_init();
    (void)s_slice_len;
    for (uint64_t i = 0; i < len; i++) {
        uint64_t phys = view_index(i, len, rotate, offset, whole_len);
        uint64_t origin = (s_begin + i) % s_len;
        fe v = load(x + 8 * phys), p = load(powers + 8 * (origin % power_num));
        fmul(&FP[0], &v, &v, &p);
        store(x + 8 * phys, &v);
    }
}

/* ---------------------------------------------------------------- curve */
typedef struct { fe x, y; } aff;
typedef struct { fe x, y, zz, zzz; } xyzz;

#define FQ (&FP[1])

static void b_mont(fe *b) {
    fe three = {{3, 0, 0, 0}};
    fmul(FQ, b, &three, &FQ->r2);
}

static inline int aff_is_id(const aff *p) { return is_zero(&p->y); }
static inline int xyzz_is_id(const xyzz *p) { return is_zero(&p->zz); }
static xyzz xyzz_identity(void) {
    xyzz r;
    memset(&r, 0, sizeof r);
    r.zzz = FQ->r1;
    return r;
}
static xyzz aff_to_xyzz(const aff *p) {
    if (aff_is_id(p)) return xyzz_identity();
    xyzz r = {p->x, p->y, FQ->r1, FQ->r1};
    return r;
}

/* mdbl-2008-s-1 (curve_xyzz.cuh:233-244) */
static xyzz aff_dbl(const aff *p) {
    fe u, v, w, s, x2, m, x3, y3, t;
    fadd(FQ, &u, &p->y, &p->y);
    fmul(FQ, &v, &u, &u);
    fmul(FQ, &w, &u, &v);
    fmul(FQ, &s, &p->x, &v);
    fmul(FQ, &x2, &p->x, &p->x);
    fadd(FQ, &m, &x2, &x2);
    fadd(FQ, &m, &m, &x2);
    fmul(FQ, &x3, &m, &m);
    fsub(FQ, &x3, &x3, &s);
    fsub(FQ, &x3, &x3, &s);
    fsub(FQ, &t, &s, &x3);
    fmul(FQ, &y3, &m, &t);
    fmul(FQ, &t, &w, &p->y);
    fsub(FQ, &y3, &y3, &t);
    xyzz r = {x3, y3, v, w};
    return r;
}

/* dbl-2008-s-1 (curve_xyzz.cuh:377-391) */
static xyzz xyzz_dbl(const xyzz *p) {
    if (xyzz_is_id(p)) return *p;
    fe u, v, w, s, x2, m, x3, y3, t, zz3, zzz3;
    fadd(FQ, &u, &p->y, &p->y);
    fmul(FQ, &v, &u, &u);
    fmul(FQ, &w, &u, &v);
    fmul(FQ, &s, &p->x, &v);
    fmul(FQ, &x2, &p->x, &p->x);
    fadd(FQ, &m, &x2, &x2);
    fadd(FQ, &m, &m, &x2);
    fmul(FQ, &x3, &m, &m);
    fsub(FQ, &x3, &x3, &s);
    fsub(FQ, &x3, &x3, &s);
    fsub(FQ, &t, &s, &x3);
    fmul(FQ, &y3, &m, &t);
    fmul(FQ, &t, &w, &p->y);
    fsub(FQ, &y3, &y3, &t);
    fmul(FQ, &zz3, &v, &p->zz);
    fmul(FQ, &zzz3, &w, &p->zzz);
    xyzz r = {x3, y3, zz3, zzz3};
    return r;
}

/* madd-2008-s (curve_xyzz.cuh:352-370) */
static xyzz xyzz_madd(const xyzz *p, const aff *q) {
    if (xyzz_is_id(p)) return aff_to_xyzz(q);
    if (aff_is_id(q)) return *p;
    fe u2, s2, pp_, r, pp, ppp, qq, x3, y3, t, zz3, zzz3;
    fmul(FQ, &u2, &q->x, &p->zz);
    fmul(FQ, &s2, &q->y, &p->zzz);
    fsub(FQ, &pp_, &u2, &p->x);
    fsub(FQ, &r, &s2, &p->y);
    if (is_zero(&pp_) && is_zero(&r)) return aff_dbl(q);
    fmul(FQ, &pp, &pp_, &pp_);
    fmul(FQ, &ppp, &pp_, &pp);
    fmul(FQ, &qq, &p->x, &pp);
    fmul(FQ, &x3, &r, &r);
    fsub(FQ, &x3, &x3, &ppp);
    fsub(FQ, &x3, &x3, &qq);
    fsub(FQ, &x3, &x3, &qq);
    fsub(FQ, &t, &qq, &x3);
    fmul(FQ, &y3, &r, &t);
    fmul(FQ, &t, &p->y, &ppp);
    fsub(FQ, &y3, &y3, &t);
    fmul(FQ, &zz3, &p->zz, &pp);
    fmul(FQ, &zzz3, &p->zzz, &ppp);
    xyzz res = {x3, y3, zz3, zzz3};
    return res;
}

/* add-2008-s (curve_xyzz.cuh:325-345) */
static xyzz xyzz_add(const xyzz *p, const xyzz *q) {
    if (xyzz_is_id(p)) return *q;
    if (xyzz_is_id(q)) return *p;
    fe u1, u2, s1, s2, pp_, r, pp, ppp, qq, x3, y3, t, zz3, zzz3;
    fmul(FQ, &u1, &p->x, &q->zz);
    fmul(FQ, &u2, &q->x, &p->zz);
    fmul(FQ, &s1, &p->y, &q->zzz);
    fmul(FQ, &s2, &q->y, &p->zzz);
    fsub(FQ, &pp_, &u2, &u1);
    fsub(FQ, &r, &s2, &s1);
    if (is_zero(&pp_) && is_zero(&r)) return xyzz_dbl(p);
    fmul(FQ, &pp, &pp_, &pp_);
    fmul(FQ, &ppp, &pp_, &pp);
    fmul(FQ, &qq, &u1, &pp);
    fmul(FQ, &x3, &r, &r);
    fsub(FQ, &x3, &x3, &ppp);
    fsub(FQ, &x3, &x3, &qq);
    fsub(FQ, &x3, &x3, &qq);
    fsub(FQ, &t, &qq, &x3);
    fmul(FQ, &y3, &r, &t);
    fmul(FQ, &t, &s1, &ppp);
    fsub(FQ, &y3, &y3, &t);
    fmul(FQ, &zz3, &p->zz, &q->zz);
    fmul(FQ, &zz3, &zz3, &pp);
    fmul(FQ, &zzz3, &p->zzz, &q->zzz);
    fmul(FQ, &zzz3, &zzz3, &ppp);
    xyzz res = {x3, y3, zz3, zzz3};
    return res;
}

/* scaling-z (curve_xyzz.cuh:316-322); identity -> (0, 0) */
static aff xyzz_to_aff(const xyzz *p) {
    aff r;
    if (xyzz_is_id(p)) {
        memset(&r, 0, sizeof r);
        return r;
    }
    fe a, b;
    finv(FQ, &a, &p->zzz);
    fmul(FQ, &b, &p->zz, &a);
    fmul(FQ, &b, &b, &b);
    fmul(FQ, &r.x, &p->x, &b);
    fmul(FQ, &r.y, &p->y, &a);
    return r;
}

static aff load_aff(const uint32_t *p) {
    aff r;
    r.x = load(p);
    r.y = load(p + 8);
    return r;
}
static void store_aff(uint32_t *p, const aff *a) {
    store(p, &a->x);
    store(p + 8, &a->y);
}
static xyzz load_xyzz(const uint32_t *p) {
    xyzz r;
    r.x = load(p);
    r.y = load(p + 8);
    r.zz = load(p + 16);
    r.zzz = load(p + 24);
    return r;
}
static void store_xyzz(uint32_t *p, const xyzz *a) {
    store(p, &a->x);
    store(p + 8, &a->y);
    store(p + 16, &a->zz);
    store(p + 24, &a->zzz);
}

void orc_xyzz_madd(uint32_t *r, const uint32_t *p, const uint32_t *q) {
    params_init();
    xyzz a = load_xyzz(p);
    aff b = load_aff(q);
    xyzz c = xyzz_madd(&a, &b);
    store_xyzz(r, &c);
}
void orc_xyzz_add(uint32_t *r, const uint32_t *p, const uint32_t *q) {
    params_init();
    xyzz a = load_xyzz(p), b = load_xyzz(q);
    xyzz c = xyzz_add(&a, &b);
    store_xyzz(r, &c);
}
void orc_xyzz_dbl(uint32_t *r, const uint32_t *p) {
    params_init();
    xyzz a = load_xyzz(p);
    xyzz c = xyzz_dbl(&a);
    store_xyzz(r, &c);
}
void orc_xyzz_to_affine(uint32_t *r, const uint32_t *p) {
    params_init();
    xyzz a = load_xyzz(p);
    aff c = xyzz_to_aff(&a);
    store_aff(r, &c);
}
void orc_affine_to_xyzz(uint32_t *r, const uint32_t *p) {
    params_init();
    aff a = load_aff(p);
    xyzz c = aff_to_xyzz(&a);
    store_xyzz(r, &c);
}
/* y^2 == x^3 + b (curve_xyzz.cuh:212-221); identity is not on the curve */
int orc_affine_on_curve(const uint32_t *p) {
    params_init();
    aff a = load_aff(p);
    fe l, rr, b;
    b_mont(&b);
    fmul(FQ, &l, &a.y, &a.y);
    fmul(FQ, &rr, &a.x, &a.x);
    fmul(FQ, &rr, &rr, &a.x);
    fadd(FQ, &rr, &rr, &b);
    return fe_eq(&l, &rr);
}
/* Y^2 == X^3 + b ZZ^3 and ZZ^3 == ZZZ^2 (curve_xyzz.cuh:296-313) */
int orc_xyzz_on_curve(const uint32_t *p) {
    params_init();
    xyzz a = load_xyzz(p);
    fe y2, x3, zz2, zz3, zzz2, b, t;
    b_mont(&b);
    fmul(FQ, &y2, &a.y, &a.y);
    fmul(FQ, &x3, &a.x, &a.x);
    fmul(FQ, &x3, &x3, &a.x);
    fmul(FQ, &zz2, &a.zz, &a.zz);
    fmul(FQ, &zz3, &zz2, &a.zz);
    fmul(FQ, &zzz2, &a.zzz, &a.zzz);
    if (!fe_eq(&zz3, &zzz2)) return 0;
    fmul(FQ, &t, &b, &zz3);
    fadd(FQ, &t, &t, &x3);
    return fe_eq(&y2, &t);
}

static aff g1_gen(void) {
    aff g;
    fe one = {{1, 0, 0, 0}}, two = {{2, 0, 0, 0}};
    fmul(FQ, &g.x, &one, &FQ->r2);
    fmul(FQ, &g.y, &two, &FQ->r2);
    return g;
}
void orc_g1_generator(uint32_t *r) {
    params_init();
    aff g = g1_gen();
    store_aff(r, &g);
}

static xyzz g1_mul_xyzz(const aff *p, const fe *k) {
    xyzz acc = xyzz_identity();
    for (int i = 255; i >= 0; i--) {
        acc = xyzz_dbl(&acc);
        if ((k->v[i >> 6] >> (i & 63)) & 1) acc = xyzz_madd(&acc, p);
    }
    return acc;
}
void orc_g1_mul(uint32_t *r, const uint32_t *p, const uint32_t *k_int) {
    params_init();
    aff a = load_aff(p);
    fe k = load(k_int);
    xyzz acc = g1_mul_xyzz(&a, &k);
    aff res = xyzz_to_aff(&acc);
    store_aff(r, &res);
}

/* fixed-base comb table: T[w][d] = d * 2^(8w) * G, w < 32, d < 256 (affine) */
static aff *comb_table(void) {
    static aff *T = NULL;
    static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    pthread_mutex_lock(&mu);
    if (!T) {
        aff *t = (aff *)calloc(32 * 256, sizeof(aff));
        aff base = g1_gen();
        for (int w = 0; w < 32; w++) {
            xyzz acc = xyzz_identity();
            for (int d = 1; d < 256; d++) {
                acc = xyzz_madd(&acc, &base);
                t[w * 256 + d] = xyzz_to_aff(&acc);
            }
            /* next base = 256 * base = 255 * base + base */
            acc = xyzz_madd(&acc, &base);
            base = xyzz_to_aff(&acc);
        }
        T = t;
    }
    pthread_mutex_unlock(&mu);
    return T;
}

typedef struct {
    uint64_t seed, start, n;
    uint32_t *out;
} bases_job;
static void *gen_bases_worker(void *arg) {
    bases_job *j = (bases_job *)arg;
    aff *T = comb_table();
    for (uint64_t t = 0; t < j->n; t++) {
        fe k = gen_uniform(j->seed, j->start + t), kint, one = {{1, 0, 0, 0}};
        fmul(&FP[0], &kint, &k, &one);
        const uint8_t *bytes = (const uint8_t *)kint.v;
        xyzz acc = xyzz_identity();
        for (int w = 0; w < 32; w++)
            if (bytes[w]) acc = xyzz_madd(&acc, &T[w * 256 + bytes[w]]);
        aff a = xyzz_to_aff(&acc);
        store_aff(j->out + 16 * t, &a);
    }
    return NULL;
}
void orc_gen_bases(uint64_t seed, uint64_t start, uint64_t n, uint32_t *out, int threads) {
    params_init();
    comb_table();
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > n) threads = n ? (int)n : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    bases_job *jobs = (bases_job *)malloc(sizeof(bases_job) * threads);
    uint64_t per = (n + threads - 1) / threads;
    int started = 0;
    for (int i = 0; i < threads; i++) {
        uint64_t lo = (uint64_t)i * per;
        if (lo >= n) break;
        uint64_t cnt = (lo + per <= n) ? per : n - lo;
        jobs[i].seed = seed;
        jobs[i].start = start + lo;
        jobs[i].n = cnt;
        jobs[i].out = out + 16 * lo;
        pthread_create(&th[i], NULL, gen_bases_worker, &jobs[i]);
        started++;
    }
    for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
    free(th);
    free(jobs);
}

/* ------------------------------------------------------------------ MSM */
/* get_at (msm_halo2.rs:74-92): c-bit digit of the little-endian repr */
static uint64_t get_at(uint64_t segment, uint64_t c, const fe *repr) {
    uint64_t skip_bits = segment * c, skip_bytes = skip_bits / 8;
    if (skip_bytes >= (254 + 7) / 8) return 0;
    const uint8_t *b = (const uint8_t *)repr->v;
    uint8_t v[8] = {0};
    for (uint64_t i = 0; i < 8 && skip_bytes + i < 32; i++) v[i] = b[skip_bytes + i];
    uint64_t tmp;
    memcpy(&tmp, v, 8);
    tmp >>= skip_bits - skip_bytes * 8;
    tmp %= (1ull << c);
    return tmp;
}

/* bucket states None / Affine / Projective (msm_halo2.rs:103-131) */
typedef struct {
    int state;
    aff a;
    xyzz p;
} bucket;

static void multiexp_serial(const fe *coeffs_mont, const aff *bases, uint64_t n, xyzz *acc) {
    uint64_t c;
    if (n < 4) c = 1;
    else if (n < 32) c = 3;
    else c = (uint64_t)ceil(log((double)(uint32_t)n)); /* f64::ln().ceil() */
    fe *repr = (fe *)malloc(sizeof(fe) * (n ? n : 1));
    fe one = {{1, 0, 0, 0}};
    for (uint64_t i = 0; i < n; i++) fmul(&FP[0], &repr[i], &coeffs_mont[i], &one);
    uint64_t segments = 254 / c + 1;
    uint64_t nb = (1ull << c) - 1;
    bucket *buckets = (bucket *)malloc(sizeof(bucket) * nb);
    for (uint64_t seg = segments; seg-- > 0;) {
        for (uint64_t i = 0; i < c; i++) *acc = xyzz_dbl(acc);
        for (uint64_t b = 0; b < nb; b++) buckets[b].state = 0;
        for (uint64_t i = 0; i < n; i++) {
            uint64_t d = get_at(seg, c, &repr[i]);
            if (d == 0) continue;
            bucket *bk = &buckets[d - 1];
            if (bk->state == 0) {
                bk->a = bases[i];
                bk->state = 1;
            } else if (bk->state == 1) {
                xyzz t = aff_to_xyzz(&bk->a);
                bk->p = xyzz_madd(&t, &bases[i]);
                bk->state = 2;
            } else {
                bk->p = xyzz_madd(&bk->p, &bases[i]);
            }
        }
        /* summation by parts (msm_halo2.rs:141-149) */
        xyzz running = xyzz_identity();
        for (uint64_t b = nb; b-- > 0;) {
            if (buckets[b].state == 1) running = xyzz_madd(&running, &buckets[b].a);
            else if (buckets[b].state == 2) running = xyzz_add(&running, &buckets[b].p);
            *acc = xyzz_add(acc, &running);
        }
    }
    free(buckets);
    free(repr);
}

typedef struct {
    const fe *coeffs;
    const aff *bases;
    uint64_t n;
    xyzz acc;
} msm_job;
static void *msm_worker(void *arg) {
    msm_job *j = (msm_job *)arg;
    j->acc = xyzz_identity();
    multiexp_serial(j->coeffs, j->bases, j->n, &j->acc);
    return NULL;
}

/* best_multiexp (msm_halo2.rs:182-209): chunk = n / threads, one serial
 * multiexp per chunk, partial results summed */
void orc_best_multiexp(const uint32_t *scalars, const uint32_t *bases_aff, uint64_t n,
                       int threads, uint32_t *r_aff) {
    params_init();
    const fe *coeffs = (const fe *)scalars;
    const aff *bases = (const aff *)bases_aff;
    xyzz total = xyzz_identity();
    if (threads < 1) threads = 1;
    if (n > (uint64_t)threads) {
        uint64_t chunk = n / threads;
        uint64_t nchunks = (n + chunk - 1) / chunk;
        msm_job *jobs = (msm_job *)malloc(sizeof(msm_job) * nchunks);
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nchunks);
        for (uint64_t i = 0; i < nchunks; i++) {
            uint64_t lo = i * chunk, cnt = (lo + chunk <= n) ? chunk : n - lo;
            jobs[i].coeffs = coeffs + lo;
            jobs[i].bases = bases + lo;
            jobs[i].n = cnt;
            pthread_create(&th[i], NULL, msm_worker, &jobs[i]);
        }
        for (uint64_t i = 0; i < nchunks; i++) {
            pthread_join(th[i], NULL);
            total = xyzz_add(&total, &jobs[i].acc);
        }
        free(jobs);
        free(th);
    } else {
        multiexp_serial(coeffs, bases, n, &total);
    }
    aff r = xyzz_to_aff(&total);
    store_aff(r_aff, &r);
}

void orc_msm_closed_form(const uint32_t *scalars, const uint32_t *k, uint64_t n, uint32_t *r_aff) {
    params_init();
    const fparams *P = &FP[0];
    fe acc = {{0, 0, 0, 0}};
    for (uint64_t i = 0; i < n; i++) {
        fe s = load(scalars + 8 * i), kk = load(k + 8 * i), t;
        fmul(P, &t, &s, &kk);
        fadd(P, &acc, &acc, &t);
    }
    fe one = {{1, 0, 0, 0}}, e;
    fmul(P, &e, &acc, &one);
    aff g = g1_gen();
    xyzz r = g1_mul_xyzz(&g, &e);
    aff a = xyzz_to_aff(&r);
    store_aff(r_aff, &a);
}

/* ----------------------------------------------------------------- gate */
void orc_gate(uint32_t *out, const uint32_t *a, const uint32_t *b, const uint32_t *c,
              const uint32_t *d, const uint32_t *q, const uint32_t *e, uint64_t n) {
    params_init();
    const fparams *P = &FP[0];
    for (uint64_t i = 0; i < n; i++) {
        fe va = load(a + 8 * i), vb = load(b + 8 * i), vc = load(c + 8 * i), vd = load(d + 8 * i),
           vq = load(q + 8 * i), ve = load(e + 8 * i), t1, t2, t3;
        fmul(P, &t1, &va, &vb);
        fmul(P, &t1, &t1, &vc);
        fsub(P, &t1, &t1, &vd);
        fmul(P, &t2, &vq, &ve);
        fsub(P, &t3, &P->r1, &vq);
        fmul(P, &t2, &t2, &t3);
        fadd(P, &t1, &t1, &t2);
        store(out + 8 * i, &t1);
    }
}

/* ------------------------------------------------- polynomial library
 * core/src/poly/src/poly.h:6-30.  Each result below is unique given its
 * inputs (field arithmetic is exact), so the restatements follow the
 * published CPU algorithms the reference's tests compare against
 * (core/tests/poly.rs:29-470): halo2_proofs kate_division and Horner
 * evaluation, ff::BatchInvert, and the tests' own prefix-product and
 * lookup-permutation truths. */

/* poly_basic.cu:4-54: r[i] = a[i] op b[i] for i < min(na, nb); past it the
 * longer operand alone, negated when it is the subtrahend.  op 0 add,
 * 1 sub, 2 mul. */
void orc_poly_binary(int op, uint32_t *r, const uint32_t *a, uint64_t na, const uint32_t *b,
                     uint64_t nb) {
    params_init();
    const fparams *P = &FP[0];
    uint64_t len = na < nb ? na : nb, upper = na < nb ? nb : na;
    for (uint64_t i = 0; i < upper; i++) {
        fe z;
        if (i < len) {
            fe x = load(a + 8 * i), y = load(b + 8 * i);
            if (op == 0) fadd(P, &z, &x, &y);
            else if (op == 1) fsub(P, &z, &x, &y);
            else fmul(P, &z, &x, &y);
        } else if (na >= nb) {
            z = load(a + 8 * i);
        } else {
            fe y = load(b + 8 * i);
            if (op == 1) fneg(P, &z, &y);
            else z = y;
        }
        store(r + 8 * i, &z);
    }
}

/* halo2_proofs::arithmetic::kate_division (called at core/tests/poly.rs:244):
 * with b' = -b, walking the coefficients from the top, q_i = a_(i+1) - tmp,
 * tmp = q_i * b'.  The n - 1 quotient coefficients go to q[0..n-1) and
 * q[n-1] = 0, the reference kernel's length (kate_division.cuh:48-88). */
void orc_kate_division(const uint32_t *p, uint64_t n, const uint32_t *b, uint32_t *q) {
    params_init();
    const fparams *P = &FP[0];
    if (n == 0) return;
    fe nb, tmp = {{0, 0, 0, 0}}, bb = load(b);
    fneg(P, &nb, &bb);
    fe zero = {{0, 0, 0, 0}};
    store(q + 8 * (n - 1), &zero);
    for (uint64_t i = n - 1; i-- > 0;) {
        fe lead = load(p + 8 * (i + 1));
        fsub(P, &lead, &lead, &tmp);
        store(q + 8 * i, &lead);
        fmul(P, &tmp, &lead, &nb);
    }
}

/* ff::BatchInvert::batch_invert (ff 0.13.1, used at core/tests/poly.rs:466):
 * Montgomery's trick over the non-zero elements; zeros stay zero. */
void orc_batch_invert(uint32_t *x, uint64_t n) {
    params_init();
    const fparams *P = &FP[0];
    fe *tmp = (fe *)malloc(sizeof(fe) * (n ? n : 1));
    fe acc = P->r1;
    for (uint64_t i = 0; i < n; i++) {
        fe v = load(x + 8 * i);
        tmp[i] = acc;
        if (!is_zero(&v)) fmul(P, &acc, &acc, &v);
    }
    finv(P, &acc, &acc);
    for (uint64_t i = n; i-- > 0;) {
        fe v = load(x + 8 * i);
        if (is_zero(&v)) continue;
        fe t;
        fmul(P, &t, &tmp[i], &acc);
        fmul(P, &acc, &acc, &v);
        store(x + 8 * i, &t);
    }
    free(tmp);
}

/* scan_mul.cuh:6-26: inclusive prefix products in place (the truth of
 * core/tests/poly.rs:412-415 after the wrapper's shift) */
void orc_scan_mul(uint32_t *x, uint64_t n) {
    params_init();
    const fparams *P = &FP[0];
    if (n == 0) return;
    fe acc = load(x);
    for (uint64_t i = 1; i < n; i++) {
        fe v = load(x + 8 * i);
        fmul(P, &acc, &acc, &v);
        store(x + 8 * i, &acc);
    }
}

static int cmp_num(const void *pa, const void *pb) {
    const fe *a = (const fe *)pa, *b = (const fe *)pb;
    for (int k = 3; k >= 0; k--) {
        if (a->v[k] != b->v[k]) return a->v[k] < b->v[k] ? -1 : 1;
    }
    return 0;
}

/* The lookup permutation of core/tests/poly.rs:556-640 (the truth the
 * reference's permute kernel is checked against, permute.cuh): the first
 * `usable` inputs sorted ascending by canonical value; at the first row of
 * each run of equal inputs the table takes that value (one copy removed
 * from the sorted table multiset), and the remaining rows take the leftover
 * table values in ascending order.  Returns the number of input runs whose
 * value is missing from the table (such rows then take the input value, no
 * copy removed); 0 for a valid lookup. */
uint64_t orc_permute(const uint32_t *input, const uint32_t *table, uint64_t usable,
                     uint32_t *res_input, uint32_t *res_table) {
    params_init();
    const fparams *P = &FP[0];
    fe one = {{1, 0, 0, 0}};
    fe *in = (fe *)malloc(sizeof(fe) * (usable ? usable : 1));
    fe *tb = (fe *)malloc(sizeof(fe) * (usable ? usable : 1));
    unsigned char *used = (unsigned char *)calloc(usable ? usable : 1, 1);
    fe *out_t = (fe *)malloc(sizeof(fe) * (usable ? usable : 1));
    for (uint64_t i = 0; i < usable; i++) {
        fe a = load(input + 8 * i), t = load(table + 8 * i);
        fmul(P, &in[i], &a, &one); /* to_number */
        fmul(P, &tb[i], &t, &one);
    }
    qsort(in, usable, sizeof(fe), cmp_num);
    qsort(tb, usable, sizeof(fe), cmp_num);
    uint64_t missing = 0, j = 0;
    for (uint64_t i = 0; i < usable; i++) {
        if (i > 0 && cmp_num(&in[i], &in[i - 1]) == 0) continue;
        while (j < usable && cmp_num(&tb[j], &in[i]) < 0) j++;
        out_t[i] = in[i];
        if (j < usable && cmp_num(&tb[j], &in[i]) == 0) {
            /* first unused copy of the value (earlier copies serve earlier runs of
             * the same value, which cannot exist: runs are distinct values) */
            used[j] = 1;
        } else {
            missing++;
        }
    }
    uint64_t k = 0;
    for (uint64_t i = 0; i < usable; i++) {
        if (i == 0 || cmp_num(&in[i], &in[i - 1]) != 0) continue;
        while (k < usable && used[k]) k++;
        out_t[i] = k < usable ? tb[k] : in[i];
        k++;
    }
    fe r2 = P->r2;
    for (uint64_t i = 0; i < usable; i++) {
        fe a, t;
        fmul(P, &a, &in[i], &r2); /* from_number */
        fmul(P, &t, &out_t[i], &r2);
        store(res_input + 8 * i, &a);
        store(res_table + 8 * i, &t);
    }
    free(in);
    free(tb);
    free(used);
    free(out_t);
    return missing;
}
