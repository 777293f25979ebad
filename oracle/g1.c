/*
 * Oracle: BLS12-381 G1 (y^2 = x^3 + 4 over Fp).  TEST INFRASTRUCTURE ONLY.
 *
 * Semantics follow vendor/midnight-curves-0.3.1/src/bls12_381/g1.rs (blst
 * p1 / p1_affine): Jacobian (X, Y, Z) with x = X/Z^2, y = Y/Z^3; Z = 0 is
 * the identity; the affine identity is (0, 0).  Group law: the textbook
 * Jacobian formulas (Cohen, Miyaji, Ono 1998), written for clarity rather
 * than speed; the accelerated code uses different (XYZZ) coordinates so a
 * shared mistake is unlikely.
 *
 * Serialisation: ZCash BLS12-381 encoding (g1.rs to_compressed /
 * to_uncompressed; params.rs:170-262 writes SRS points uncompressed).
 *
 * Pinned by: 2G in Montgomery limbs (g1.rs:1589-1622), [r]G = O, on-curve
 * checks, the reference SRS file transient-crypto/static/bls_midnight_2p14
 * (g[0] is the generator, every point decodes on-curve), and Python
 * affine big-integer arithmetic (tests/test_oracle_g1.py).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

/* canonical generator coordinates (big-endian hex in the BLS12-381 spec) */
static const uint64_t GX[6] = {
    0xfb3af00adb22c6bbULL, 0x6c55e83ff97a1aefULL, 0xa14e3a3f171bac58ULL,
    0xc3688c4f9774b905ULL, 0x2695638c4fa9ac0fULL, 0x17f1d3a73197d794ULL};
static const uint64_t GY[6] = {
    0x0caa232946c5e7e1ULL, 0xd03cc744a2888ae4ULL, 0x00db18cb2c04b3edULL,
    0xfcf5e095d5d00af6ULL, 0xa09e30ed741d8ae4ULL, 0x08b3f481e3aaa0f1ULL};
/* (p - 1) / 2, canonical */
static const uint64_t HALF_P[6] = {
    0xdcff7fffffffd555ULL, 0x0f55ffff58a9ffffULL, 0xb39869507b587b12ULL,
    0xb23ba5c279c2895fULL, 0x258dd3db21a5d66bULL, 0x0d0088f51cbff34dULL};

void orc_g1_generator(orc_g1a *r)
{
    orc_fp_from_canonical(&r->x, GX);
    orc_fp_from_canonical(&r->y, GY);
}

void orc_g1j_identity(orc_g1j *r)
{
    memset(r, 0, sizeof(*r));
    orc_fp_one(&r->x);
    orc_fp_one(&r->y);
}

int orc_g1j_is_identity(const orc_g1j *p) { return orc_fp_is_zero(&p->z); }
int orc_g1a_is_identity(const orc_g1a *p) { return orc_fp_is_zero(&p->x) && orc_fp_is_zero(&p->y); }

void orc_g1j_from_affine(orc_g1j *r, const orc_g1a *p)
{
    if (orc_g1a_is_identity(p)) { orc_g1j_identity(r); return; }
    r->x = p->x;
    r->y = p->y;
    orc_fp_one(&r->z);
}

void orc_g1j_to_affine(orc_g1a *r, const orc_g1j *p)
{
    if (orc_g1j_is_identity(p)) { memset(r, 0, sizeof(*r)); return; }
    orc_fp zi, zi2, zi3;
    orc_fp_inv(&zi, &p->z);
    orc_fp_sqr(&zi2, &zi);
    orc_fp_mul(&zi3, &zi2, &zi);
    orc_fp_mul(&r->x, &p->x, &zi2);
    orc_fp_mul(&r->y, &p->y, &zi3);
}

void orc_g1_batch_to_affine(orc_g1a *out, const orc_g1j *in, size_t n)
{
    /* Montgomery's simultaneous inversion over the non-identity inputs */
    if (n == 0) return;
    orc_fp *prefix = (orc_fp *)malloc(n * sizeof(orc_fp));
    orc_fp acc;
    orc_fp_one(&acc);
    for (size_t i = 0; i < n; ++i) {
        prefix[i] = acc;
        if (!orc_g1j_is_identity(&in[i])) orc_fp_mul(&acc, &acc, &in[i].z);
    }
    orc_fp inv;
    orc_fp_inv(&inv, &acc);
    for (size_t i = n; i-- > 0;) {
        if (orc_g1j_is_identity(&in[i])) { memset(&out[i], 0, sizeof(out[i])); continue; }
        orc_fp zi, zi2, zi3;
        orc_fp_mul(&zi, &inv, &prefix[i]);
        orc_fp_mul(&inv, &inv, &in[i].z);
        orc_fp_sqr(&zi2, &zi);
        orc_fp_mul(&zi3, &zi2, &zi);
        orc_fp_mul(&out[i].x, &in[i].x, &zi2);
        orc_fp_mul(&out[i].y, &in[i].y, &zi3);
    }
    free(prefix);
}

void orc_g1j_dbl(orc_g1j *r, const orc_g1j *p)
{
    if (orc_g1j_is_identity(p) || orc_fp_is_zero(&p->y)) { orc_g1j_identity(r); return; }
    orc_fp yy, s, m, t, x3, y3, z3, yyyy;
    orc_fp_sqr(&yy, &p->y);
    orc_fp_mul(&s, &p->x, &yy);
    orc_fp_add(&s, &s, &s);
    orc_fp_add(&s, &s, &s);                 /* S = 4 X Y^2 */
    orc_fp_sqr(&m, &p->x);
    orc_fp_add(&t, &m, &m);
    orc_fp_add(&m, &t, &m);                 /* M = 3 X^2 */
    orc_fp_sqr(&x3, &m);
    orc_fp_sub(&x3, &x3, &s);
    orc_fp_sub(&x3, &x3, &s);               /* X3 = M^2 - 2S */
    orc_fp_sqr(&yyyy, &yy);
    orc_fp_add(&yyyy, &yyyy, &yyyy);
    orc_fp_add(&yyyy, &yyyy, &yyyy);
    orc_fp_add(&yyyy, &yyyy, &yyyy);        /* 8 Y^4 */
    orc_fp_sub(&t, &s, &x3);
    orc_fp_mul(&y3, &m, &t);
    orc_fp_sub(&y3, &y3, &yyyy);            /* Y3 = M (S - X3) - 8 Y^4 */
    orc_fp_mul(&z3, &p->y, &p->z);
    orc_fp_add(&z3, &z3, &z3);              /* Z3 = 2 Y Z */
    r->x = x3; r->y = y3; r->z = z3;
}

void orc_g1j_add(orc_g1j *r, const orc_g1j *p, const orc_g1j *q)
{
    if (orc_g1j_is_identity(p)) { *r = *q; return; }
    if (orc_g1j_is_identity(q)) { *r = *p; return; }
    orc_fp z1z1, z2z2, u1, u2, s1, s2, h, rr, hh, hhh, v, x3, y3, z3, t;
    orc_fp_sqr(&z1z1, &p->z);
    orc_fp_sqr(&z2z2, &q->z);
    orc_fp_mul(&u1, &p->x, &z2z2);
    orc_fp_mul(&u2, &q->x, &z1z1);
    orc_fp_mul(&s1, &p->y, &z2z2);
    orc_fp_mul(&s1, &s1, &q->z);
    orc_fp_mul(&s2, &q->y, &z1z1);
    orc_fp_mul(&s2, &s2, &p->z);
    orc_fp_sub(&h, &u2, &u1);
    orc_fp_sub(&rr, &s2, &s1);
    if (orc_fp_is_zero(&h)) {
        if (orc_fp_is_zero(&rr)) { orc_g1j_dbl(r, p); return; }
        orc_g1j_identity(r);
        return;
    }
    orc_fp_sqr(&hh, &h);
    orc_fp_mul(&hhh, &hh, &h);
    orc_fp_mul(&v, &u1, &hh);
    orc_fp_sqr(&x3, &rr);
    orc_fp_sub(&x3, &x3, &hhh);
    orc_fp_sub(&x3, &x3, &v);
    orc_fp_sub(&x3, &x3, &v);               /* X3 = R^2 - H^3 - 2 U1 H^2 */
    orc_fp_sub(&t, &v, &x3);
    orc_fp_mul(&y3, &rr, &t);
    orc_fp_mul(&t, &s1, &hhh);
    orc_fp_sub(&y3, &y3, &t);               /* Y3 = R (U1 H^2 - X3) - S1 H^3 */
    orc_fp_mul(&z3, &p->z, &q->z);
    orc_fp_mul(&z3, &z3, &h);               /* Z3 = H Z1 Z2 */
    r->x = x3; r->y = y3; r->z = z3;
}

void orc_g1j_add_affine(orc_g1j *r, const orc_g1j *p, const orc_g1a *q)
{
    orc_g1j qj;
    orc_g1j_from_affine(&qj, q);
    orc_g1j_add(r, p, &qj);
}

void orc_g1j_neg(orc_g1j *r, const orc_g1j *p)
{
    *r = *p;
    orc_fp_neg(&r->y, &p->y);
}

int orc_g1j_eq(const orc_g1j *p, const orc_g1j *q)
{
    int ip = orc_g1j_is_identity(p), iq = orc_g1j_is_identity(q);
    if (ip || iq) return ip && iq;
    orc_fp z1z1, z2z2, a, b;
    orc_fp_sqr(&z1z1, &p->z);
    orc_fp_sqr(&z2z2, &q->z);
    orc_fp_mul(&a, &p->x, &z2z2);
    orc_fp_mul(&b, &q->x, &z1z1);
    if (!orc_fp_eq(&a, &b)) return 0;
    orc_fp_mul(&a, &p->y, &z2z2);
    orc_fp_mul(&a, &a, &q->z);
    orc_fp_mul(&b, &q->y, &z1z1);
    orc_fp_mul(&b, &b, &p->z);
    return orc_fp_eq(&a, &b);
}

int orc_g1a_is_on_curve(const orc_g1a *p)
{
    if (orc_g1a_is_identity(p)) return 1;
    orc_fp y2, x3, four, one;
    orc_fp_sqr(&y2, &p->y);
    orc_fp_sqr(&x3, &p->x);
    orc_fp_mul(&x3, &x3, &p->x);
    orc_fp_one(&one);
    orc_fp_add(&four, &one, &one);
    orc_fp_add(&four, &four, &four);
    orc_fp_add(&x3, &x3, &four);
    return orc_fp_eq(&y2, &x3);
}

void orc_g1j_mul(orc_g1j *r, const orc_g1j *p, const orc_fr *s)
{
    uint64_t e[4];
    orc_fr_to_canonical(e, s);
    orc_g1j acc;
    orc_g1j_identity(&acc);
    for (int i = 255; i >= 0; --i) {
        orc_g1j_dbl(&acc, &acc);
        if ((e[i / 64] >> (i % 64)) & 1) orc_g1j_add(&acc, &acc, p);
    }
    *r = acc;
}

void orc_g1_mul(orc_g1j *r, const orc_g1a *p, const orc_fr *s)
{
    orc_g1j pj;
    orc_g1j_from_affine(&pj, p);
    orc_g1j_mul(r, &pj, s);
}

/* ---------------- ZCash encoding ---------------- */

static void fp_to_be(uint8_t out[48], const orc_fp *a)
{
    uint64_t c[6];
    orc_fp_to_canonical(c, a);
    for (int i = 0; i < 6; ++i)
        for (int b = 0; b < 8; ++b)
            out[47 - (8 * i + b)] = (uint8_t)(c[i] >> (8 * b));
}

static int fp_from_be(orc_fp *r, const uint8_t in[48])
{
    uint64_t c[6] = {0};
    for (int i = 0; i < 6; ++i)
        for (int b = 0; b < 8; ++b)
            c[i] |= (uint64_t)in[47 - (8 * i + b)] << (8 * b);
    return orc_fp_from_canonical(r, c);
}

static int fp_lexicographically_largest(const orc_fp *y)
{
    uint64_t c[6];
    orc_fp_to_canonical(c, y);
    for (int i = 5; i >= 0; --i) {
        if (c[i] > HALF_P[i]) return 1;
        if (c[i] < HALF_P[i]) return 0;
    }
    return 0;
}

void orc_g1_compress(uint8_t out[48], const orc_g1a *p)
{
    if (orc_g1a_is_identity(p)) {
        memset(out, 0, 48);
        out[0] = 0xc0;
        return;
    }
    fp_to_be(out, &p->x);
    out[0] |= 0x80;
    if (fp_lexicographically_largest(&p->y)) out[0] |= 0x20;
}

int orc_g1_decompress(orc_g1a *r, const uint8_t in[48])
{
    uint8_t buf[48];
    memcpy(buf, in, 48);
    int compressed = (buf[0] >> 7) & 1, inf = (buf[0] >> 6) & 1, sign = (buf[0] >> 5) & 1;
    if (!compressed) return 0;
    buf[0] &= 0x1f;
    if (inf) {
        for (int i = 0; i < 48; ++i) if (buf[i]) return 0;
        if (sign) return 0;
        memset(r, 0, sizeof(*r));
        return 1;
    }
    orc_fp x, rhs, y, one, four;
    if (!fp_from_be(&x, buf)) return 0;
    orc_fp_sqr(&rhs, &x);
    orc_fp_mul(&rhs, &rhs, &x);
    orc_fp_one(&one);
    orc_fp_add(&four, &one, &one);
    orc_fp_add(&four, &four, &four);
    orc_fp_add(&rhs, &rhs, &four);
    if (!orc_fp_sqrt(&y, &rhs)) return 0;
    if (fp_lexicographically_largest(&y) != sign) orc_fp_neg(&y, &y);
    r->x = x;
    r->y = y;
    return 1;
}

void orc_g1_to_uncompressed(uint8_t out[96], const orc_g1a *p)
{
    if (orc_g1a_is_identity(p)) {
        memset(out, 0, 96);
        out[0] = 0x40;
        return;
    }
    fp_to_be(out, &p->x);
    fp_to_be(out + 48, &p->y);
}

int orc_g1_from_uncompressed(orc_g1a *r, const uint8_t in[96])
{
    uint8_t buf[96];
    memcpy(buf, in, 96);
    int compressed = (buf[0] >> 7) & 1, inf = (buf[0] >> 6) & 1, sign = (buf[0] >> 5) & 1;
    if (compressed || sign) return 0;
    buf[0] &= 0x1f;
    if (inf) {
        for (int i = 0; i < 96; ++i) if (buf[i]) return 0;
        memset(r, 0, sizeof(*r));
        return 1;
    }
    if (!fp_from_be(&r->x, buf)) return 0;
    if (!fp_from_be(&r->y, buf + 48)) return 0;
    return orc_g1a_is_on_curve(r);
}

/* decompress n consecutive 48-byte points; returns the number that failed */
size_t orc_g1_decompress_many(orc_g1a *out, const uint8_t *in, size_t n, int threads)
{
    size_t bad = 0;
    if (threads < 1) threads = 1;
    #pragma omp parallel for num_threads(threads) reduction(+ : bad) schedule(static)
    for (long long i = 0; i < (long long)n; ++i)
        if (!orc_g1_decompress(&out[i], in + 48 * i)) ++bad;
    return bad;
}
