/*
 * Oracle: G1 multi-scalar multiplication and SRS helpers.
 * TEST INFRASTRUCTURE ONLY.
 *
 * orc_msm_pippenger restates vendor/midnight-curves-0.3.1/src/msm.rs:477-560
 * (msm_best): window width c = ceil(ln n), NUM_BITS / c + 1 windows, Booth
 * digits from get_booth_index (msm.rs:13-50), 2^(c-1) buckets per window,
 * summation by parts, c*w doublings per window, sum over windows.  The
 * reference's batch-affine bucket schedule (msm.rs:52-430) changes only the
 * order of additions; here every bucket add is a Jacobian+affine add.
 * Windows (and, when threads outnumber windows, slices of the points) are
 * spread over OpenMP threads, as msm_best spreads windows over rayon.
 *
 * The MSM value is unique, so any correct algorithm gives the same point.
 * Pinned by: orc_msm_naive on small n (independent double-and-add), the
 * closed form MSM([t_i]G, s) = [sum s_i t_i]G, the KZG closed form
 * MSM([tau^i]G, c) = [c(tau)]G, and the reference SRS file identity
 * MSM(g, c) == MSM(g_lagrange, NTT(c)) (tests/test_oracle_msm.py).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* mixed addition (madd-2007-bl), handling identity / doubling / inverse */
static void jadd_aff(orc_g1j *r, const orc_g1j *p, const orc_g1a *q, int negate)
{
    if (orc_g1a_is_identity(q)) { *r = *p; return; }
    orc_fp qy = q->y;
    if (negate) orc_fp_neg(&qy, &q->y);
    if (orc_g1j_is_identity(p)) {
        r->x = q->x; r->y = qy; orc_fp_one(&r->z);
        return;
    }
    orc_fp z1z1, u2, s2, h, hh, i4, j, rr, v, t, x3, y3, z3;
    orc_fp_sqr(&z1z1, &p->z);
    orc_fp_mul(&u2, &q->x, &z1z1);
    orc_fp_mul(&s2, &qy, &p->z);
    orc_fp_mul(&s2, &s2, &z1z1);
    orc_fp_sub(&h, &u2, &p->x);
    orc_fp_sub(&rr, &s2, &p->y);
    if (orc_fp_is_zero(&h)) {
        if (orc_fp_is_zero(&rr)) { orc_g1j_dbl(r, p); return; }
        orc_g1j_identity(r);
        return;
    }
    orc_fp_add(&rr, &rr, &rr);              /* r = 2 (S2 - Y1) */
    orc_fp_sqr(&hh, &h);
    orc_fp_add(&i4, &hh, &hh);
    orc_fp_add(&i4, &i4, &i4);              /* I = 4 HH */
    orc_fp_mul(&j, &h, &i4);                /* J = H I */
    orc_fp_mul(&v, &p->x, &i4);             /* V = X1 I */
    orc_fp_sqr(&x3, &rr);
    orc_fp_sub(&x3, &x3, &j);
    orc_fp_sub(&x3, &x3, &v);
    orc_fp_sub(&x3, &x3, &v);               /* X3 = r^2 - J - 2V */
    orc_fp_sub(&t, &v, &x3);
    orc_fp_mul(&y3, &rr, &t);
    orc_fp_mul(&t, &p->y, &j);
    orc_fp_add(&t, &t, &t);
    orc_fp_sub(&y3, &y3, &t);               /* Y3 = r (V - X3) - 2 Y1 J */
    orc_fp_add(&z3, &p->z, &h);
    orc_fp_sqr(&z3, &z3);
    orc_fp_sub(&z3, &z3, &z1z1);
    orc_fp_sub(&z3, &z3, &hh);              /* Z3 = (Z1 + H)^2 - Z1Z1 - HH */
    r->x = x3; r->y = y3; r->z = z3;
}

void orc_msm_naive(orc_g1j *out, const orc_g1a *pts, const orc_fr *sc, size_t n)
{
    orc_g1j acc, t;
    orc_g1j_identity(&acc);
    for (size_t i = 0; i < n; ++i) {
        orc_g1_mul(&t, &pts[i], &sc[i]);
        orc_g1j_add(&acc, &acc, &t);
    }
    *out = acc;
}

int32_t orc_booth_index(size_t window_index, size_t window_size, const uint8_t el[32])
{
    size_t skip_bits = window_index * window_size;
    skip_bits = skip_bits ? skip_bits - 1 : 0;
    size_t skip_bytes = skip_bits / 8;
    uint8_t v[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < 4 && skip_bytes + i < 32; ++i) v[i] = el[skip_bytes + i];
    uint32_t tmp = (uint32_t)v[0] | (uint32_t)v[1] << 8 | (uint32_t)v[2] << 16 | (uint32_t)v[3] << 24;
    if (window_index == 0) tmp <<= 1;
    tmp >>= skip_bits - skip_bytes * 8;
    tmp &= (1u << (window_size + 1)) - 1;
    int sign = (tmp & (1u << window_size)) == 0;
    tmp = (tmp + 1) >> 1;
    if (sign) return (int32_t)tmp;
    return -(int32_t)((~(tmp - 1)) & ((1u << window_size) - 1));
}

void orc_msm_pippenger(orc_g1j *out, const orc_g1a *pts, const orc_fr *sc, size_t n, int threads)
{
    if (threads < 1) threads = 1;
    orc_g1j_identity(out);
    if (n == 0) return;
    if (n < 32) { orc_msm_naive(out, pts, sc, n); return; }
    size_t c = (size_t)ceil(log((double)n));
    size_t nwin = 255 / c + 1;
    size_t nb = (size_t)1 << (c - 1);

    uint8_t *bytes = (uint8_t *)malloc(n * 32);
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        uint64_t can[4];
        orc_fr_to_canonical(can, &sc[i]);
        for (int l = 0; l < 4; ++l)
            for (int b = 0; b < 8; ++b) bytes[32 * i + 8 * l + b] = (uint8_t)(can[l] >> (8 * b));
    }

    /* split points into slices when threads outnumber windows */
    size_t slices = (size_t)threads > nwin ? ((size_t)threads + nwin - 1) / nwin : 1;
    if (slices > n / 1024 + 1) slices = n / 1024 + 1;
    size_t tasks = nwin * slices;
    orc_g1j *part = (orc_g1j *)malloc(tasks * sizeof(orc_g1j));

    #pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (long long task = 0; task < (long long)tasks; ++task) {
        size_t w = (size_t)task / slices, sl = (size_t)task % slices;
        size_t lo = n * sl / slices, hi = n * (sl + 1) / slices;
        orc_g1j *buck = (orc_g1j *)malloc(nb * sizeof(orc_g1j));
        for (size_t b = 0; b < nb; ++b) orc_g1j_identity(&buck[b]);
        for (size_t i = lo; i < hi; ++i) {
            int32_t d = orc_booth_index(w, c, bytes + 32 * i);
            if (d == 0) continue;
            size_t b = (size_t)(d > 0 ? d : -d) - 1;
            jadd_aff(&buck[b], &buck[b], &pts[i], d < 0);
        }
        orc_g1j run, acc;
        orc_g1j_identity(&run);
        orc_g1j_identity(&acc);
        for (size_t b = nb; b-- > 0;) {
            orc_g1j_add(&run, &run, &buck[b]);
            orc_g1j_add(&acc, &acc, &run);
        }
        for (size_t k = 0; k < c * w; ++k) orc_g1j_dbl(&acc, &acc);
        part[task] = acc;
        free(buck);
    }
    orc_g1j total;
    orc_g1j_identity(&total);
    for (size_t t = 0; t < tasks; ++t) orc_g1j_add(&total, &total, &part[t]);
    *out = total;
    free(part);
    free(bytes);
}

/* ---------------------------------------------------------------------
 * orc_msm_best: vendor/midnight-curves-0.3.1/src/msm.rs:477-560 msm_best
 * with its batch-affine bucket schedule (msm.rs:52-430), the CPU baseline's
 * algorithm.  Window c = ceil(ln n) (n < 8104 gives c < 10, where msm_best
 * takes msm_parallel: orc_msm_pippenger here); identity bases filtered;
 * NUM_BITS / c + 1 windows spread over threads; per window, each nonzero
 * Booth digit either goes to an affine bucket through a schedule of up to
 * BATCH = 64 pending additions (a bucket that is still empty takes the point
 * outright; a full schedule runs batch_add: one inversion for all its
 * additions, Montgomery's trick) or, when its bucket already has a pending
 * addition in the schedule, into a Jacobian bucket; then summation by parts
 * over Jacobian + affine buckets and c * w doublings.
 */
#define MSMB_BATCH 64

typedef struct { size_t base, buck; int sign; } msmb_sched;

/* batch_add (msm.rs:58-142) */
static void msmb_batch_add(orc_g1a *ab, uint8_t *ainf, const msmb_sched *set, size_t size, const orc_g1a *bases)
{
    orc_fp t[MSMB_BATCH], z[MSMB_BATCH], acc;
    orc_fp_one(&acc);
    for (size_t i = 0; i < size; ++i) {
        const size_t b = set[i].buck;
        const orc_g1a *p = &bases[set[i].base];
        if (ainf[b]) continue;
        if (orc_fp_eq(&ab[b].x, &p->x)) {
            if (orc_fp_eq(&ab[b].y, &p->y) ^ !set[i].sign) {
                /* doubling: lambda = 3 x^2 / 2 y */
                orc_fp x2, x3;
                orc_fp_sqr(&x2, &p->x);
                orc_fp_add(&x3, &x2, &x2);
                orc_fp_add(&x3, &x3, &x2);
                orc_fp_add(&z[i], &ab[b].y, &ab[b].y);
                orc_fp_mul(&t[i], &acc, &x3);
                orc_fp_mul(&acc, &acc, &z[i]);
                continue;
            }
            ainf[b] = 1;   /* P + (-P) */
            continue;
        }
        orc_fp dy;
        orc_fp_sub(&z[i], &ab[b].x, &p->x);
        if (set[i].sign) orc_fp_sub(&dy, &ab[b].y, &p->y);
        else orc_fp_add(&dy, &ab[b].y, &p->y);
        orc_fp_mul(&t[i], &acc, &dy);
        orc_fp_mul(&acc, &acc, &z[i]);
    }
    orc_fp_inv_vartime(&acc, &acc);
    for (size_t i = size; i-- > 0;) {
        const size_t b = set[i].buck;
        const orc_g1a *p = &bases[set[i].base];
        if (ainf[b]) continue;
        orc_fp lambda, x, y, d;
        orc_fp_mul(&lambda, &acc, &t[i]);
        orc_fp_mul(&acc, &acc, &z[i]);
        orc_fp_sqr(&x, &lambda);
        orc_fp_add(&d, &ab[b].x, &p->x);
        orc_fp_sub(&x, &x, &d);
        orc_fp_sub(&d, &p->x, &x);
        orc_fp_mul(&y, &lambda, &d);
        if (set[i].sign) orc_fp_sub(&y, &y, &p->y);
        else orc_fp_add(&y, &y, &p->y);
        ab[b].x = x;
        ab[b].y = y;
    }
}

void orc_msm_best(orc_g1j *out, const orc_g1a *pts, const orc_fr *sc, size_t n, int threads)
{
    if (threads < 1) threads = 1;
    orc_g1j_identity(out);
    if (n == 0) return;
    size_t c = n < 4 ? 1 : n < 32 ? 3 : (size_t)ceil(log((double)n));
    if (c < 10) { orc_msm_pippenger(out, pts, sc, n, threads); return; }
    const size_t nb = (size_t)1 << (c - 1);
    const size_t nwin = 255 / c + 1;

    /* filter identities; scalars to little-endian bytes */
    size_t *idx = (size_t *)malloc(n * sizeof(size_t));
    size_t m = 0;
    for (size_t i = 0; i < n; ++i)
        if (!orc_g1a_is_identity(&pts[i])) idx[m++] = i;
    uint8_t *bytes = (uint8_t *)malloc(m * 32);
    orc_g1a *bases = (orc_g1a *)malloc(m * sizeof(orc_g1a));
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (long long k = 0; k < (long long)m; ++k) {
        uint64_t can[4];
        orc_fr_to_canonical(can, &sc[idx[k]]);
        for (int l = 0; l < 4; ++l)
            for (int b = 0; b < 8; ++b) bytes[32 * k + 8 * l + b] = (uint8_t)(can[l] >> (8 * b));
        bases[k] = pts[idx[k]];
    }
    orc_g1j *acc = (orc_g1j *)malloc(nwin * sizeof(orc_g1j));

    #pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (long long wi = (long long)nwin - 1; wi >= 0; --wi) {
        const size_t w = (size_t)wi;
        orc_g1j *jb = (orc_g1j *)malloc(nb * sizeof(orc_g1j));
        orc_g1a *ab = (orc_g1a *)malloc(nb * sizeof(orc_g1a));
        uint8_t *ainf = (uint8_t *)malloc(nb);
        for (size_t b = 0; b < nb; ++b) orc_g1j_identity(&jb[b]);
        memset(ainf, 1, nb);
        msmb_sched set[MSMB_BATCH];
        size_t ptr = 0;
        for (size_t k = 0; k < m; ++k) {
            const int32_t d = orc_booth_index(w, c, bytes + 32 * k);
            if (d == 0) continue;
            const int sign = d > 0;
            const size_t b = (size_t)(d > 0 ? d : -d) - 1;
            int pending = 0;
            for (size_t q = 0; q < ptr; ++q)
                if (set[q].buck == b) { pending = 1; break; }
            if (pending) {
                jadd_aff(&jb[b], &jb[b], &bases[k], !sign);   /* greedy Jacobian accumulation */
                continue;
            }
            if (ainf[b]) {
                ab[b] = bases[k];
                if (!sign) orc_fp_neg(&ab[b].y, &ab[b].y);
                ainf[b] = 0;
            } else {
                set[ptr].base = k;
                set[ptr].buck = b;
                set[ptr].sign = sign;
                if (++ptr == MSMB_BATCH) {
                    msmb_batch_add(ab, ainf, set, ptr, bases);
                    ptr = 0;
                }
            }
        }
        if (ptr) msmb_batch_add(ab, ainf, set, ptr, bases);
        orc_g1j run, a, s;
        orc_g1j_identity(&run);
        orc_g1j_identity(&a);
        for (size_t b = nb; b-- > 0;) {
            s = jb[b];
            if (!ainf[b]) jadd_aff(&s, &s, &ab[b], 0);
            orc_g1j_add(&run, &run, &s);
            orc_g1j_add(&a, &a, &run);
        }
        for (size_t k = 0; k < c * w; ++k) orc_g1j_dbl(&a, &a);
        acc[w] = a;
        free(jb);
        free(ab);
        free(ainf);
    }
    orc_g1j total;
    orc_g1j_identity(&total);
    for (size_t w = 0; w < nwin; ++w) orc_g1j_add(&total, &total, &acc[w]);
    *out = total;
    free(acc);
    free(bases);
    free(bytes);
    free(idx);
}

/* fixed-base comb: table[j][b] = [b * 256^j] G, j < 32, b < 256 */
static orc_g1a *g_table;

static void build_table(void)
{
    if (g_table) return;
    orc_g1j *tj = (orc_g1j *)malloc(32 * 256 * sizeof(orc_g1j));
    orc_g1a g;
    orc_g1_generator(&g);
    orc_g1j base;
    orc_g1j_from_affine(&base, &g);
    for (int j = 0; j < 32; ++j) {
        orc_g1j_identity(&tj[j * 256]);
        for (int b = 1; b < 256; ++b) orc_g1j_add(&tj[j * 256 + b], &tj[j * 256 + b - 1], &base);
        for (int k = 0; k < 8; ++k) orc_g1j_dbl(&base, &base);
    }
    orc_g1a *t = (orc_g1a *)malloc(32 * 256 * sizeof(orc_g1a));
    orc_g1_batch_to_affine(t, tj, 32 * 256);
    free(tj);
    g_table = t;
}

static void fixed_base(orc_g1j *r, const orc_fr *s)
{
    uint64_t e[4];
    orc_fr_to_canonical(e, s);
    orc_g1j acc;
    orc_g1j_identity(&acc);
    for (int j = 0; j < 32; ++j) {
        unsigned b = (unsigned)(e[j / 8] >> (8 * (j % 8))) & 0xff;
        if (b) jadd_aff(&acc, &acc, &g_table[j * 256 + b], 0);
    }
    *r = acc;
}

void orc_g1_fixed_base_batch(orc_g1a *out, const orc_fr *t, size_t n, int threads)
{
    if (threads < 1) threads = 1;
    #pragma omp critical(orc_table)
    build_table();
    const size_t B = 4096;
    #pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (long long blk = 0; blk < (long long)((n + B - 1) / B); ++blk) {
        size_t lo = (size_t)blk * B, hi = lo + B < n ? lo + B : n;
        orc_g1j *tmp = (orc_g1j *)malloc((hi - lo) * sizeof(orc_g1j));
        for (size_t i = lo; i < hi; ++i) fixed_base(&tmp[i - lo], &t[i]);
        orc_g1_batch_to_affine(out + lo, tmp, hi - lo);
        free(tmp);
    }
}

void orc_srs_powers(orc_g1a *out, const orc_fr *tau, size_t n, int threads)
{
    orc_fr *pw = (orc_fr *)malloc(n * sizeof(orc_fr));
    if (n) orc_fr_one(&pw[0]);
    for (size_t i = 1; i < n; ++i) orc_fr_mul(&pw[i], &pw[i - 1], tau);
    orc_g1_fixed_base_batch(out, pw, n, threads);
    free(pw);
}
