/*
 * Coefficient-form polynomial helpers of the KZG opening -- CPU oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates vendor/midnight-proofs-0.8.1/src/utils/arithmetic.rs:
 *   eval_polynomial     :53-78   (Horner from the top coefficient)
 *   kate_division_owned :119-135 (quotient by X - b, remainder dropped)
 * Pinned by tests/test_oracle_prover.py: evaluation against a direct power
 * sum in Python integers, and q(X) (X - b) + a(b) = a(X) coefficient-wise.
 */
#include "oracle.h"
#include <string.h>

void orc_poly_eval(orc_fr *out, const orc_fr *poly, size_t n, const orc_fr *point)
{
    orc_fr acc; memset(&acc, 0, sizeof acc);
    for (size_t i = n; i-- > 0;) {
        orc_fr_mul(&acc, &acc, point);
        orc_fr_add(&acc, &acc, &poly[i]);
    }
    *out = acc;
}

/* q[0 .. n-2] = a / (X - b); q[n-1] = 0 (q may alias a) */
void orc_kate_division(orc_fr *q, const orc_fr *a, size_t n, const orc_fr *b)
{
    if (n == 0) return;
    orc_fr nb; orc_fr_neg(&nb, b);
    orc_fr tmp; memset(&tmp, 0, sizeof tmp);
    orc_fr lead;
    /* top-down: lead = a[idx] - tmp, tmp = lead * (-b) */
    orc_fr *w = q;
    if (q != a) memcpy(w, a, n * sizeof(orc_fr));
    for (size_t idx = n - 1; idx >= 1; --idx) {
        orc_fr_sub(&lead, &w[idx], &tmp);
        w[idx] = lead;
        orc_fr_mul(&tmp, &lead, &nb);
    }
    memmove(w, w + 1, (n - 1) * sizeof(orc_fr));
    memset(&w[n - 1], 0, sizeof(orc_fr));
}
