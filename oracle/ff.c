/*
 * Oracle: Montgomery arithmetic for BLS12-381 Fr (4 x u64, R = 2^256) and
 * Fp (6 x u64, R = 2^384).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows the semantics of vendor/midnight-curves-0.3.1/src/bls12_381/fq.rs
 * (Fr; blst_fr) and fp.rs (Fp; blst_fp): values are kept fully reduced in
 * [0, modulus) and in Montgomery form, so equal field elements have equal
 * limbs.  The product is textbook CIOS (Koc, Acar, Kaliski 1996).
 *
 * Pinned by: the reference's limb KATs (fq.rs:1147-1175 mul,
 * fp.rs:1258-1290 mul, the R / R^2 / R^3 constants fq.rs:88-128), Fermat
 * inverse round trips, and Python big-integer arithmetic on random inputs
 * (tests/test_oracle_field.py).
 */
#include "oracle.h"
#include <string.h>

typedef unsigned __int128 u128;

/* ---------------- generic n-limb helpers (n <= 6) ---------------- */

static inline int geq_n(const uint64_t *a, const uint64_t *b, int n)
{
    for (int i = n - 1; i >= 0; --i) {
        if (a[i] > b[i]) return 1;
        if (a[i] < b[i]) return 0;
    }
    return 1;
}

static inline uint64_t sub_n(uint64_t *r, const uint64_t *a, const uint64_t *b, int n)
{
    uint64_t borrow = 0;
    for (int i = 0; i < n; ++i) {
        u128 d = (u128)a[i] - b[i] - borrow;
        r[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) & 1;
    }
    return borrow;
}

static inline uint64_t add_n(uint64_t *r, const uint64_t *a, const uint64_t *b, int n)
{
    uint64_t carry = 0;
    for (int i = 0; i < n; ++i) {
        u128 s = (u128)a[i] + b[i] + carry;
        r[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
    return carry;
}

static inline void mod_add(uint64_t *r, const uint64_t *a, const uint64_t *b,
                           const uint64_t *p, int n)
{
    uint64_t t[6];
    uint64_t c = add_n(t, a, b, n);
    if (c || geq_n(t, p, n)) sub_n(t, t, p, n);
    memcpy(r, t, 8 * n);
}

static inline void mod_sub(uint64_t *r, const uint64_t *a, const uint64_t *b,
                           const uint64_t *p, int n)
{
    uint64_t t[6];
    if (sub_n(t, a, b, n)) add_n(t, t, p, n);
    memcpy(r, t, 8 * n);
}

/* CIOS Montgomery product r = a*b/2^(64n) mod p, inputs < p */
static inline void mont_mul(uint64_t *r, const uint64_t *a, const uint64_t *b,
                            const uint64_t *p, uint64_t inv, int n)
{
    uint64_t t[8] = {0};
    for (int i = 0; i < n; ++i) {
        u128 c = 0;
        for (int j = 0; j < n; ++j) {
            c = (u128)a[j] * b[i] + t[j] + (uint64_t)(c >> 64);
            t[j] = (uint64_t)c;
        }
        u128 s = (u128)t[n] + (uint64_t)(c >> 64);
        t[n] = (uint64_t)s;
        t[n + 1] = (uint64_t)(s >> 64);
        uint64_t m = t[0] * inv;
        c = (u128)m * p[0] + t[0];
        for (int j = 1; j < n; ++j) {
            c = (u128)m * p[j] + t[j] + (uint64_t)(c >> 64);
            t[j - 1] = (uint64_t)c;
        }
        s = (u128)t[n] + (uint64_t)(c >> 64);
        t[n - 1] = (uint64_t)s;
        t[n] = t[n + 1] + (uint64_t)(s >> 64);
    }
    if (t[n] || geq_n(t, p, n)) sub_n(t, t, p, n);
    memcpy(r, t, 8 * n);
}

/* -(p^-1) mod 2^64 by Newton iteration */
static uint64_t neg_inv64(uint64_t p0)
{
    uint64_t x = 1;
    for (int i = 0; i < 7; ++i) x *= 2 - p0 * x;
    return (uint64_t)0 - x;
}

/* 2^bits mod p by doubling (slow; constants only) */
static void pow2_mod(uint64_t *r, unsigned bits, const uint64_t *p, int n)
{
    uint64_t t[6] = {1, 0, 0, 0, 0, 0};
    for (unsigned i = 0; i < bits; ++i) mod_add(t, t, t, p, n);
    memcpy(r, t, 8 * n);
}

/* ---------------- Fr ---------------- */

const orc_fr ORC_FR_MODULUS = {{
    0xffffffff00000001ULL, 0x53bda402fffe5bfeULL,
    0x3339d80809a1d805ULL, 0x73eda753299d7d48ULL}};

static uint64_t FR_INV;
static orc_fr FR_R, FR_R2, FR_R3;

static const uint64_t FP_P[6] = {
    0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL,
    0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};
static uint64_t FP_INV;
static orc_fp FP_R, FP_R2, FP_R3;

__attribute__((constructor)) static void orc_ff_init(void)
{
    FR_INV = neg_inv64(ORC_FR_MODULUS.l[0]);
    pow2_mod(FR_R.l, 256, ORC_FR_MODULUS.l, 4);
    pow2_mod(FR_R2.l, 512, ORC_FR_MODULUS.l, 4);
    pow2_mod(FR_R3.l, 768, ORC_FR_MODULUS.l, 4);
    FP_INV = neg_inv64(FP_P[0]);
    pow2_mod(FP_R.l, 384, FP_P, 6);
    pow2_mod(FP_R2.l, 768, FP_P, 6);
    pow2_mod(FP_R3.l, 1152, FP_P, 6);
}

void orc_fr_add(orc_fr *r, const orc_fr *a, const orc_fr *b) { mod_add(r->l, a->l, b->l, ORC_FR_MODULUS.l, 4); }
void orc_fr_sub(orc_fr *r, const orc_fr *a, const orc_fr *b) { mod_sub(r->l, a->l, b->l, ORC_FR_MODULUS.l, 4); }
void orc_fr_neg(orc_fr *r, const orc_fr *a)
{
    orc_fr z = {{0}};
    orc_fr_sub(r, &z, a);
}
void orc_fr_mul(orc_fr *r, const orc_fr *a, const orc_fr *b) { mont_mul(r->l, a->l, b->l, ORC_FR_MODULUS.l, FR_INV, 4); }
void orc_fr_sqr(orc_fr *r, const orc_fr *a) { orc_fr_mul(r, a, a); }
void orc_fr_one(orc_fr *r) { *r = FR_R; }
int orc_fr_eq(const orc_fr *a, const orc_fr *b) { return memcmp(a, b, sizeof(*a)) == 0; }
int orc_fr_is_zero(const orc_fr *a) { return (a->l[0] | a->l[1] | a->l[2] | a->l[3]) == 0; }

void orc_fr_pow(orc_fr *r, const orc_fr *a, const uint64_t e[4])
{
    orc_fr acc = FR_R, base = *a;
    for (int i = 0; i < 256; ++i) {
        if ((e[i / 64] >> (i % 64)) & 1) orc_fr_mul(&acc, &acc, &base);
        orc_fr_sqr(&base, &base);
    }
    *r = acc;
}

void orc_fr_inv(orc_fr *r, const orc_fr *a)
{
    uint64_t e[4], two[4] = {2, 0, 0, 0};
    sub_n(e, ORC_FR_MODULUS.l, two, 4);   /* Fermat: a^(r-2) */
    orc_fr_pow(r, a, e);
}

void orc_fr_from_u64(orc_fr *r, uint64_t v)
{
    orc_fr t = {{v, 0, 0, 0}};
    orc_fr_mul(r, &t, &FR_R2);
}

void orc_fr_to_canonical(uint64_t out[4], const orc_fr *a)
{
    orc_fr one = {{1, 0, 0, 0}}, t;
    orc_fr_mul(&t, a, &one);
    memcpy(out, t.l, 32);
}

int orc_fr_from_canonical(orc_fr *r, const uint64_t in[4])
{
    orc_fr t;
    memcpy(t.l, in, 32);
    if (geq_n(t.l, ORC_FR_MODULUS.l, 4)) return 0;
    orc_fr_mul(r, &t, &FR_R2);
    return 1;
}

/* Fq::from_uniform_bytes: split 64 bytes into two 256-bit limbs a0, a1
 * taken as *Montgomery-form* values (not reduced), result a0*R2 + a1*R3.
 * The Montgomery product tolerates inputs < 2^256 because r < 2^255 and
 * the CIOS bound t < 2r holds for a*b < r*2^256. */
void orc_fr_from_uniform_bytes(orc_fr *r, const uint8_t bytes[64])
{
    orc_fr a0, a1, t0, t1;
    memcpy(a0.l, bytes, 32);
    memcpy(a1.l, bytes + 32, 32);
    orc_fr_mul(&t0, &a0, &FR_R2);
    orc_fr_mul(&t1, &a1, &FR_R3);
    orc_fr_add(r, &t0, &t1);
}

void orc_fr_root_of_unity(orc_fr *r)
{
    static const orc_fr v = {{0xb9b58d8c5f0e466aULL, 0x5b1b4c801819d7ecULL,
                              0x0af53ae352a31e64ULL, 0x5bf3adda19e9b27bULL}};
    *r = v;
}

void orc_fr_zeta(orc_fr *r)
{
    static const orc_fr v = {{0x92d9090b093011d2ULL, 0xfc9cbd719d6aa073ULL,
                              0xc1f14ef0cd65a1a6ULL, 0x017f6d35e72fcdebULL}};
    *r = v;
}

void orc_fr_delta(orc_fr *r)
{
    static const orc_fr v = {{0x70e310d3d146f96aULL, 0x4b64c08919e299e6ULL,
                              0x51e114186a8b970dULL, 0x6185d06627c067cbULL}};
    *r = v;
}

/* ---------------- Fp ---------------- */

void orc_fp_add(orc_fp *r, const orc_fp *a, const orc_fp *b) { mod_add(r->l, a->l, b->l, FP_P, 6); }
void orc_fp_sub(orc_fp *r, const orc_fp *a, const orc_fp *b) { mod_sub(r->l, a->l, b->l, FP_P, 6); }
void orc_fp_neg(orc_fp *r, const orc_fp *a)
{
    orc_fp z = {{0}};
    orc_fp_sub(r, &z, a);
}
void orc_fp_mul(orc_fp *r, const orc_fp *a, const orc_fp *b) { mont_mul(r->l, a->l, b->l, FP_P, FP_INV, 6); }
void orc_fp_sqr(orc_fp *r, const orc_fp *a) { orc_fp_mul(r, a, a); }
void orc_fp_one(orc_fp *r) { *r = FP_R; }
int orc_fp_eq(const orc_fp *a, const orc_fp *b) { return memcmp(a, b, sizeof(*a)) == 0; }
int orc_fp_is_zero(const orc_fp *a)
{
    uint64_t o = 0;
    for (int i = 0; i < 6; ++i) o |= a->l[i];
    return o == 0;
}

static void fp_pow(orc_fp *r, const orc_fp *a, const uint64_t e[6])
{
    orc_fp acc = FP_R, base = *a;
    for (int i = 0; i < 384; ++i) {
        if ((e[i / 64] >> (i % 64)) & 1) orc_fp_mul(&acc, &acc, &base);
        orc_fp_sqr(&base, &base);
    }
    *r = acc;
}

/* variable-time inverse by the binary extended Euclidean algorithm (for
 * the CPU baseline's batch inversions; orc_fp_inv pins it): x = aR as an
 * integer, x^-1 = a^-1 R^-1, and a Montgomery product with R^3 gives a^-1 R */
void orc_fp_inv_vartime(orc_fp *r, const orc_fp *a)
{
    if (orc_fp_is_zero(a)) { memset(r, 0, sizeof(*r)); return; }
    uint64_t u[6], v[6], x1[6] = {1, 0, 0, 0, 0, 0}, x2[6] = {0};
    memcpy(u, a->l, 48);
    memcpy(v, FP_P, 48);
    const uint64_t one[6] = {1, 0, 0, 0, 0, 0};
    for (;;) {
        if (memcmp(u, one, 48) == 0 || memcmp(v, one, 48) == 0) break;
        while (!(u[0] & 1)) {
            for (int i = 0; i < 5; ++i) u[i] = (u[i] >> 1) | (u[i + 1] << 63);
            u[5] >>= 1;
            uint64_t c = 0;
            if (x1[0] & 1) c = add_n(x1, x1, FP_P, 6);
            for (int i = 0; i < 5; ++i) x1[i] = (x1[i] >> 1) | (x1[i + 1] << 63);
            x1[5] = (x1[5] >> 1) | (c << 63);
        }
        while (!(v[0] & 1)) {
            for (int i = 0; i < 5; ++i) v[i] = (v[i] >> 1) | (v[i + 1] << 63);
            v[5] >>= 1;
            uint64_t c = 0;
            if (x2[0] & 1) c = add_n(x2, x2, FP_P, 6);
            for (int i = 0; i < 5; ++i) x2[i] = (x2[i] >> 1) | (x2[i + 1] << 63);
            x2[5] = (x2[5] >> 1) | (c << 63);
        }
        if (geq_n(u, v, 6)) {
            sub_n(u, u, v, 6);
            mod_sub(x1, x1, x2, FP_P, 6);
        } else {
            sub_n(v, v, u, 6);
            mod_sub(x2, x2, x1, FP_P, 6);
        }
    }
    orc_fp t;
    memcpy(t.l, memcmp(u, one, 48) == 0 ? x1 : x2, 48);
    orc_fp_mul(r, &t, &FP_R3);
}

void orc_fp_inv(orc_fp *r, const orc_fp *a)
{
    uint64_t e[6], two[6] = {2, 0, 0, 0, 0, 0};
    sub_n(e, FP_P, two, 6);
    fp_pow(r, a, e);
}

/* p = 3 mod 4: sqrt(a) = a^((p+1)/4) */
int orc_fp_sqrt(orc_fp *r, const orc_fp *a)
{
    uint64_t e[6], one[6] = {1, 0, 0, 0, 0, 0};
    add_n(e, FP_P, one, 6);
    for (int i = 0; i < 6; ++i) e[i] = (e[i] >> 2) | (i < 5 ? e[i + 1] << 62 : 0);
    orc_fp s, s2;
    fp_pow(&s, a, e);
    orc_fp_sqr(&s2, &s);
    *r = s;
    return orc_fp_eq(&s2, a);
}

void orc_fp_to_canonical(uint64_t out[6], const orc_fp *a)
{
    orc_fp one = {{1, 0, 0, 0, 0, 0}}, t;
    orc_fp_mul(&t, a, &one);
    memcpy(out, t.l, 48);
}

int orc_fp_from_canonical(orc_fp *r, const uint64_t in[6])
{
    orc_fp t;
    memcpy(t.l, in, 48);
    if (geq_n(t.l, FP_P, 6)) return 0;
    orc_fp_mul(r, &t, &FP_R2);
    return 1;
}

/* sum_i a_i * b_i (closed form of an MSM over points [t_i]G) */
void orc_fr_inner_product(orc_fr *r, const orc_fr *a, const orc_fr *b, size_t n)
{
    orc_fr acc = {{0}}, t;
    for (size_t i = 0; i < n; ++i) {
        orc_fr_mul(&t, &a[i], &b[i]);
        orc_fr_add(&acc, &acc, &t);
    }
    *r = acc;
}

/* out[i] = a[i] * b[i] + c[i], element by element (test data builders) */
void orc_fr_mul_add_many(orc_fr *out, const orc_fr *a, const orc_fr *b, const orc_fr *c, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        orc_fr t;
        orc_fr_mul(&t, &a[i], &b[i]);
        orc_fr_add(&out[i], &t, &c[i]);
    }
}

/* out[i] = sum_j s[j] * polys[j * n + i] (m polynomials of n coefficients) */
void orc_fr_lincomb(orc_fr *out, const orc_fr *polys, const orc_fr *s, size_t m, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        orc_fr acc = {{0}}, t;
        for (size_t j = 0; j < m; ++j) {
            orc_fr_mul(&t, &s[j], &polys[j * n + i]);
            orc_fr_add(&acc, &acc, &t);
        }
        out[i] = acc;
    }
}
