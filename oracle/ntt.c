/*
 * Oracle: Fr NTT, evaluation-domain constants and coset transforms.
 * TEST INFRASTRUCTURE ONLY.
 *
 * orc_best_fft restates vendor/midnight-curves-0.3.1/src/fft.rs:30-112
 * (bit-reverse permutation, then log n radix-2 DIT stages with twiddles
 * omega^i, no scaling); the rayon recursion of fft.rs:180-216 computes the
 * same butterflies in another order and is replaced by OpenMP over the
 * blocks of each stage.  The domain functions restate
 * vendor/midnight-proofs-0.8.1/src/poly/domain.rs:54-159 (constants),
 * :325-338 (coeff_to_extended), :410-443 (extended_to_coeff),
 * :476-506 (divide_by_vanishing_poly), :544-592 (zeta distribution, ifft).
 *
 * Pinned by: the DFT definition evaluated directly on small sizes,
 * iNTT(NTT(x)) = x, NTT(x^j) = omega^(ij), zeta^3 = 1, t(X) = X^n - 1 on the
 * coset, and the reference SRS file (MSM(g, c) == MSM(g_lagrange, NTT(c)))
 * (tests/test_oracle_ntt.py).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

static size_t bitrev(size_t x, unsigned bits)
{
    size_t r = 0;
    for (unsigned i = 0; i < bits; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

void orc_best_fft(orc_fr *a, const orc_fr *omega, uint32_t log_n, int threads)
{
    size_t n = (size_t)1 << log_n;
    if (threads < 1) threads = 1;
    for (size_t k = 0; k < n; ++k) {
        size_t rk = bitrev(k, log_n);
        if (k < rk) { orc_fr t = a[k]; a[k] = a[rk]; a[rk] = t; }
    }
    if (n < 2) return;
    orc_fr *tw = (orc_fr *)malloc((n / 2) * sizeof(orc_fr));
    orc_fr_one(&tw[0]);
    for (size_t i = 1; i < n / 2; ++i) orc_fr_mul(&tw[i], &tw[i - 1], omega);

    size_t chunk = 2, tchunk = n / 2;
    for (uint32_t s = 0; s < log_n; ++s) {
        size_t half = chunk / 2, nblocks = n / chunk;
        /* flatten (block, i) so that early and late stages both parallelise */
        long long total = (long long)(n / 2);
        #pragma omp parallel for num_threads(threads) schedule(static) if (total >= 4096)
        for (long long f = 0; f < total; ++f) {
            size_t blk = (size_t)f / half, i = (size_t)f % half;
            orc_fr *lo = a + blk * chunk, *hi = lo + half;
            orc_fr t = hi[i];
            if (i) orc_fr_mul(&t, &t, &tw[i * tchunk]);
            orc_fr u = lo[i];
            orc_fr_add(&lo[i], &u, &t);
            orc_fr_sub(&hi[i], &u, &t);
        }
        (void)nblocks;
        chunk *= 2;
        tchunk /= 2;
    }
    free(tw);
}

void orc_omega(orc_fr *r, uint32_t log_n)
{
    orc_fr w;
    orc_fr_root_of_unity(&w);
    for (uint32_t i = log_n; i < 32; ++i) orc_fr_sqr(&w, &w);
    *r = w;
}

void orc_mul_scalar(orc_fr *a, size_t n, const orc_fr *s)
{
    #pragma omp parallel for schedule(static) if (n >= 65536)
    for (long long i = 0; i < (long long)n; ++i) orc_fr_mul(&a[i], &a[i], s);
}

void orc_fft(orc_fr *a, uint32_t log_n, int threads)
{
    orc_fr w;
    orc_omega(&w, log_n);
    orc_best_fft(a, &w, log_n, threads);
}

void orc_ifft(orc_fr *a, uint32_t log_n, int threads)
{
    orc_fr w, wi, ninv, nn;
    orc_omega(&w, log_n);
    orc_fr_inv(&wi, &w);
    orc_best_fft(a, &wi, log_n, threads);
    orc_fr_from_u64(&nn, (uint64_t)1 << log_n);
    orc_fr_inv(&ninv, &nn);
    orc_mul_scalar(a, (size_t)1 << log_n, &ninv);
}

void orc_mul_zeta(orc_fr *a, size_t n, const orc_fr *first, const orc_fr *second)
{
    #pragma omp parallel for schedule(static) if (n >= 65536)
    for (long long i = 0; i < (long long)n; ++i) {
        int c = (int)(i % 3);
        if (c == 1) orc_fr_mul(&a[i], &a[i], first);
        else if (c == 2) orc_fr_mul(&a[i], &a[i], second);
    }
}

void orc_mul_repeating(orc_fr *a, size_t n, const orc_fr *f, size_t m)
{
    #pragma omp parallel for schedule(static) if (n >= 65536)
    for (long long i = 0; i < (long long)n; ++i) orc_fr_mul(&a[i], &a[i], &f[(size_t)i % m]);
}

void orc_domain_new(orc_domain *d, orc_fr *t_evals, uint32_t j, uint32_t k)
{
    uint64_t n = (uint64_t)1 << k, qdeg = (uint64_t)(j - 1);
    uint32_t ek = k;
    while (((uint64_t)1 << ek) < n * qdeg) ++ek;
    d->k = k;
    d->extended_k = ek;
    d->j = j;
    orc_omega(&d->extended_omega, ek);
    d->omega = d->extended_omega;
    for (uint32_t i = k; i < ek; ++i) orc_fr_sqr(&d->omega, &d->omega);
    orc_fr_inv(&d->omega_inv, &d->omega);
    orc_fr_inv(&d->extended_omega_inv, &d->extended_omega);
    orc_fr_zeta(&d->g_coset);
    orc_fr_sqr(&d->g_coset_inv, &d->g_coset);
    orc_fr t;
    orc_fr_from_u64(&t, n);
    orc_fr_inv(&d->ifft_divisor, &t);
    orc_fr_from_u64(&t, (uint64_t)1 << ek);
    orc_fr_inv(&d->extended_ifft_divisor, &t);
    orc_fr_from_u64(&t, n);
    orc_fr_inv(&d->barycentric_weight, &t);
    if (t_evals) {
        /* t(zeta * ext_omega^i) = zeta^n * (ext_omega^n)^i - 1, inverted */
        uint64_t e[4] = {n, 0, 0, 0};
        orc_fr orig, step, cur, one;
        orc_fr_pow(&orig, &d->g_coset, e);
        orc_fr_pow(&step, &d->extended_omega, e);
        orc_fr_one(&one);
        cur = orig;
        size_t cnt = (size_t)1 << (ek - k);
        for (size_t i = 0; i < cnt; ++i) {
            orc_fr v;
            orc_fr_sub(&v, &cur, &one);
            orc_fr_inv(&t_evals[i], &v);
            orc_fr_mul(&cur, &cur, &step);
        }
    }
}

void orc_coeff_to_extended(orc_fr *out, const orc_fr *coeffs, const orc_domain *d, int threads)
{
    size_t n = (size_t)1 << d->k, N = (size_t)1 << d->extended_k;
    memcpy(out, coeffs, n * sizeof(orc_fr));
    orc_mul_zeta(out, n, &d->g_coset, &d->g_coset_inv);
    memset(out + n, 0, (N - n) * sizeof(orc_fr));
    orc_best_fft(out, &d->extended_omega, d->extended_k, threads);
}

void orc_extended_to_coeff(orc_fr *a, const orc_domain *d, int threads)
{
    size_t N = (size_t)1 << d->extended_k;
    orc_best_fft(a, &d->extended_omega_inv, d->extended_k, threads);
    orc_mul_scalar(a, N, &d->extended_ifft_divisor);
    orc_mul_zeta(a, N, &d->g_coset_inv, &d->g_coset);
}
