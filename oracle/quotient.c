/*
 * Quotient numerator on the extended domain -- CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates the reference CPU prover's numerator evaluation
 * (vendor/midnight-proofs-0.8.1/src/plonk/evaluation.rs):
 *   - ValueSource::get            evaluation.rs:84-119
 *   - Calculation::evaluate       evaluation.rs:150-210
 *   - GraphEvaluator::evaluate    evaluation.rs:3797-3845
 *   - get_rotation_idx            evaluation.rs:42-45
 *   - Evaluator::evaluate_numerator, one proof:
 *       custom gates              evaluation.rs:3190-3225
 *       permutation               evaluation.rs:3258-3355
 *       lookups                   evaluation.rs:3382-3470
 *       trashcans                 evaluation.rs:3473-3525
 * over the compiled graph encoding that the GPU entry point receives
 * (RawGraph / RawCalculation / RawValueSource, bls12_381_g1_msm.cu:351-381;
 * opcode numbers midnight-gpu-prover/src/expression.rs:80-98).
 *
 * The phases run one after another over all rows, as the CPU prover does;
 * each phase folds its terms into the running value with value = value*y + t.
 *
 * Parity pinning (tests/test_oracle_quotient.py):
 *   - the graph evaluator against a direct evaluation of the expression the
 *     graph was compiled from (gate q*(a*b + c - d), Horner in y, squares,
 *     doubles, negations, constants, rotations in both directions);
 *   - the whole numerator evaluated on the base domain of a satisfying
 *     synthetic circuit (gate + 8-column copy constraints split over
 *     3 permutation sets) is zero on every row, and a single changed witness
 *     cell makes exactly the expected rows nonzero;
 *   - the numerator on the extended coset of that circuit, divided by the
 *     vanishing polynomial, interpolates to a polynomial whose top
 *     coefficients (degree >= (n-1)(d-1)) are zero;
 *   - lookup products against Python integers on a tiny input (with a zero
 *     denominator and the blinded tail), and closure to one at the last
 *     usable row when the permuted columns permute the compressed ones.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

static int64_t rot_idx(size_t idx, int32_t rot, int32_t rot_scale, size_t len)
{
    /* get_rotation_idx: (idx + rot*rot_scale) rem_euclid size */
    int64_t m = (int64_t)len;
    int64_t s = ((int64_t)idx + (int64_t)rot * (int64_t)rot_scale) % m;
    return s < 0 ? s + m : s;
}

static const orc_fr *column_base(uint32_t kind, uint32_t index, const orc_numerator_inputs *in)
{
    if (kind == 0 && index < in->fixed_cols) return in->fixed + (size_t)index * in->len;
    if (kind == 1 && index < in->advice_cols) return in->advice + (size_t)index * in->len;
    if (kind == 2 && index < in->instance_cols) return in->instance + (size_t)index * in->len;
    return NULL;
}

static orc_fr source_value(const orc_value_source *s, const orc_graph *g, size_t idx,
                           const orc_numerator_inputs *in, const orc_fr *inter,
                           const orc_fr *previous)
{
    orc_fr z; memset(&z, 0, sizeof z);
    switch (s->kind) {
    case 0: return s->a < g->constants_len ? g->constants[s->a] : z;
    case 1: return s->a < g->num_intermediates ? inter[s->a] : z;
    case 2: case 3: case 4: {
        const orc_fr *base = column_base(s->kind - 2, s->a, in);
        if (!base || s->b >= g->rotations_len) return z;
        return base[rot_idx(idx, g->rotations[s->b], in->rot_scale, in->len)];
    }
    case 5: return s->a < in->challenge_len ? in->challenges[s->a] : z;
    case 6: return in->beta;
    case 7: return in->gamma;
    case 8: return in->theta;
    case 9: return in->trash_challenge;
    case 10: return in->y;
    case 11: return *previous;
    default: return z;
    }
}

/* GraphEvaluator::evaluate for one row; inter holds num_intermediates
 * values, zeroed here (compiled graphs never read a slot before writing
 * it in the same row). */
static orc_fr graph_eval(const orc_graph *g, size_t idx, const orc_numerator_inputs *in,
                         orc_fr *inter, const orc_fr *previous)
{
    orc_fr last; memset(&last, 0, sizeof last);
    memset(inter, 0, sizeof(orc_fr) * g->num_intermediates);
    for (size_t c = 0; c < g->calculations_len; c++) {
        const orc_calculation *calc = &g->calculations[c];
        orc_fr v, a, b;
        memset(&v, 0, sizeof v);
        switch (calc->op) {
        case 0: a = source_value(&calc->lhs, g, idx, in, inter, previous);
                b = source_value(&calc->rhs, g, idx, in, inter, previous);
                orc_fr_add(&v, &a, &b); break;
        case 1: a = source_value(&calc->lhs, g, idx, in, inter, previous);
                b = source_value(&calc->rhs, g, idx, in, inter, previous);
                orc_fr_sub(&v, &a, &b); break;
        case 2: a = source_value(&calc->lhs, g, idx, in, inter, previous);
                b = source_value(&calc->rhs, g, idx, in, inter, previous);
                orc_fr_mul(&v, &a, &b); break;
        case 3: a = source_value(&calc->lhs, g, idx, in, inter, previous);
                orc_fr_sqr(&v, &a); break;
        case 4: a = source_value(&calc->lhs, g, idx, in, inter, previous);
                orc_fr_add(&v, &a, &a); break;
        case 5: a = source_value(&calc->lhs, g, idx, in, inter, previous);
                orc_fr_neg(&v, &a); break;
        case 6: {   /* Horner(start, parts, factor) */
            orc_fr f = source_value(&calc->rhs, g, idx, in, inter, previous);
            v = source_value(&calc->lhs, g, idx, in, inter, previous);
            for (uint32_t p = 0; p < calc->parts_len; p++) {
                size_t pi = (size_t)calc->parts_offset + p;
                orc_fr part; memset(&part, 0, sizeof part);
                if (pi < g->horner_parts_len)
                    part = source_value(&g->horner_parts[pi], g, idx, in, inter, previous);
                orc_fr_mul(&v, &v, &f);
                orc_fr_add(&v, &v, &part);
            }
            break;
        }
        case 7: v = source_value(&calc->lhs, g, idx, in, inter, previous); break;
        default: break;
        }
        if (calc->target < g->num_intermediates) inter[calc->target] = v;
        last = v;
    }
    return last;
}

void orc_graph_eval_rows(orc_fr *out, const orc_graph *g, const orc_numerator_inputs *in,
                         const orc_fr *previous, int threads)
{
    size_t len = in->len;
    #pragma omp parallel num_threads(threads > 0 ? threads : 1)
    {
        orc_fr *inter = (orc_fr *)calloc(g->num_intermediates ? g->num_intermediates : 1, sizeof(orc_fr));
        orc_fr zero; memset(&zero, 0, sizeof zero);
        #pragma omp for schedule(static)
        for (size_t i = 0; i < len; i++)
            out[i] = graph_eval(g, i, in, inter, previous ? &previous[i] : &zero);
        free(inter);
    }
}

static inline void fold(orc_fr *v, const orc_fr *y, const orc_fr *t)
{
    orc_fr_mul(v, v, y);
    orc_fr_add(v, v, t);
}

void orc_eval_numerator(orc_fr *out, const orc_numerator_inputs *in, int threads)
{
    const size_t len = in->len;
    const orc_fr y = in->y, beta = in->beta, gamma = in->gamma;
    orc_fr one; orc_fr_one(&one);
    if (threads <= 0) threads = 1;

    /* custom gates (evaluation.rs:3190-3225): the graph's closing Horner
     * reads PreviousValue, the running value of the row */
    if (in->run_custom) {
        /* one proof: the running value is zero when the gates are folded */
        orc_graph_eval_rows(out, &in->custom_graph, in, NULL, threads);
    } else {
        memcpy(out, in->initial_out, len * sizeof(orc_fr));
    }

    /* permutation (evaluation.rs:3258-3355) */
    if (in->permutation_sets != 0 && in->chunk_len != 0) {
        orc_fr delta_start;
        orc_fr_mul(&delta_start, &beta, &in->g_coset);
        orc_fr delta; orc_fr_delta(&delta);
        const size_t sets = in->permutation_sets;
        const orc_fr *first = in->permutation_products;
        const orc_fr *lastp = in->permutation_products + (sets - 1) * len;
        /* parallelize() splits the rows into chunks; beta_term = omega^start
         * at a chunk start, then *= omega per row (evaluation.rs:3283-3352) */
        const size_t nchunks = (size_t)threads * 4;
        const size_t chunk = (len + nchunks - 1) / nchunks;
        #pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
        for (size_t ch = 0; ch < nchunks; ch++) {
            size_t start = ch * chunk, end = start + chunk < len ? start + chunk : len;
            orc_fr beta_term;
            uint64_t e[4] = {(uint64_t)start, 0, 0, 0};
            orc_fr_pow(&beta_term, &in->omega, e);
            for (size_t idx = start; idx < end; idx++) {
                size_t r_next = (size_t)rot_idx(idx, 1, in->rot_scale, len);
                size_t r_last = (size_t)rot_idx(idx, in->last_rotation, in->rot_scale, len);
                orc_fr v = out[idx], t, u;

                orc_fr_sub(&t, &one, &first[idx]);
                orc_fr_mul(&t, &t, &in->l0[idx]);
                fold(&v, &y, &t);

                orc_fr_sqr(&t, &lastp[idx]);
                orc_fr_sub(&t, &t, &lastp[idx]);
                orc_fr_mul(&t, &t, &in->l_last[idx]);
                fold(&v, &y, &t);

                for (size_t s = 1; s < sets; s++) {
                    orc_fr_sub(&t, &in->permutation_products[s * len + idx],
                               &in->permutation_products[(s - 1) * len + r_last]);
                    orc_fr_mul(&t, &t, &in->l0[idx]);
                    fold(&v, &y, &t);
                }

                orc_fr current_delta;
                orc_fr_mul(&current_delta, &delta_start, &beta_term);
                for (size_t s = 0; s < sets; s++) {
                    const orc_fr *z = in->permutation_products + s * len;
                    size_t base = s * in->chunk_len;
                    orc_fr left = z[r_next], right = z[idx];
                    for (size_t j = 0; j < in->chunk_len && base + j < in->permutation_columns_len; j++) {
                        const orc_column_ref *c = &in->permutation_columns[base + j];
                        const orc_fr *col = column_base(c->kind, c->index, in);
                        if (!col) continue;
                        const orc_fr *sigma = in->permutation_pk_cosets + (base + j) * len;
                        orc_fr_mul(&t, &beta, &sigma[idx]);
                        orc_fr_add(&t, &t, &col[idx]);
                        orc_fr_add(&t, &t, &gamma);
                        orc_fr_mul(&left, &left, &t);
                    }
                    for (size_t j = 0; j < in->chunk_len && base + j < in->permutation_columns_len; j++) {
                        const orc_column_ref *c = &in->permutation_columns[base + j];
                        const orc_fr *col = column_base(c->kind, c->index, in);
                        if (!col) continue;
                        orc_fr_add(&t, &col[idx], &current_delta);
                        orc_fr_add(&t, &t, &gamma);
                        orc_fr_mul(&right, &right, &t);
                        orc_fr_mul(&current_delta, &current_delta, &delta);
                    }
                    orc_fr_sub(&u, &left, &right);
                    orc_fr_mul(&u, &u, &in->l_active[idx]);
                    fold(&v, &y, &u);
                }
                out[idx] = v;
                orc_fr_mul(&beta_term, &beta_term, &in->omega);
            }
        }
    }

    /* lookups (evaluation.rs:3382-3470): all lookups of a row in order */
    if (in->lookup_count != 0) {
        const orc_fr zero = {{0, 0, 0, 0}};
        #pragma omp parallel num_threads(threads)
        {
            uint32_t maxi = 1;
            for (size_t n = 0; n < in->lookup_count; n++)
                if (in->lookup_graphs[n].num_intermediates > maxi) maxi = in->lookup_graphs[n].num_intermediates;
            orc_fr *inter = (orc_fr *)calloc(maxi, sizeof(orc_fr));
            #pragma omp for schedule(static)
            for (size_t idx = 0; idx < len; idx++) {
                size_t r_next = (size_t)rot_idx(idx, 1, in->rot_scale, len);
                size_t r_prev = (size_t)rot_idx(idx, -1, in->rot_scale, len);
                orc_fr v = out[idx], t, u, w;
                for (size_t n = 0; n < in->lookup_count; n++) {
                    const orc_fr *z = in->lookup_product_cosets + n * len;
                    const orc_fr *ai = in->lookup_input_cosets + n * len;
                    const orc_fr *st = in->lookup_table_cosets + n * len;
                    orc_fr table_value = graph_eval(&in->lookup_graphs[n], idx, in, inter, &zero);
                    orc_fr a_minus_s; orc_fr_sub(&a_minus_s, &ai[idx], &st[idx]);

                    orc_fr_sub(&t, &one, &z[idx]);
                    orc_fr_mul(&t, &t, &in->l0[idx]);
                    fold(&v, &y, &t);

                    orc_fr_sqr(&t, &z[idx]);
                    orc_fr_sub(&t, &t, &z[idx]);
                    orc_fr_mul(&t, &t, &in->l_last[idx]);
                    fold(&v, &y, &t);

                    orc_fr_add(&t, &ai[idx], &beta);
                    orc_fr_mul(&t, &z[r_next], &t);
                    orc_fr_add(&u, &st[idx], &gamma);
                    orc_fr_mul(&t, &t, &u);
                    orc_fr_mul(&w, &z[idx], &table_value);
                    orc_fr_sub(&t, &t, &w);
                    orc_fr_mul(&t, &t, &in->l_active[idx]);
                    fold(&v, &y, &t);

                    orc_fr_mul(&t, &a_minus_s, &in->l0[idx]);
                    fold(&v, &y, &t);

                    orc_fr_sub(&t, &ai[idx], &ai[r_prev]);
                    orc_fr_mul(&t, &a_minus_s, &t);
                    orc_fr_mul(&t, &t, &in->l_active[idx]);
                    fold(&v, &y, &t);
                }
                out[idx] = v;
            }
            free(inter);
        }
    }

    /* trashcans (evaluation.rs:3473-3525), one pass per trashcan */
    for (size_t n = 0; n < in->trash_count; n++) {
        const orc_graph *g = &in->trash_graphs[n];
        uint32_t sel = in->trash_selector_fixed_indices[n];
        const orc_fr *trash = in->trash_cosets + n * len;
        const orc_fr zero = {{0, 0, 0, 0}};
        if (sel >= in->fixed_cols) continue;
        #pragma omp parallel num_threads(threads)
        {
            orc_fr *inter = (orc_fr *)calloc(g->num_intermediates ? g->num_intermediates : 1, sizeof(orc_fr));
            #pragma omp for schedule(static)
            for (size_t idx = 0; idx < len; idx++) {
                orc_fr compressed = graph_eval(g, idx, in, inter, &zero);
                orc_fr q = in->fixed[(size_t)sel * len + idx], t;
                orc_fr_sub(&t, &one, &q);
                orc_fr_mul(&t, &t, &trash[idx]);
                orc_fr_sub(&t, &compressed, &t);
                fold(&out[idx], &y, &t);
            }
            free(inter);
        }
    }
}

/* Permutation grand products (permutation/prover.rs:355-515), one column
 * set of chunk_len columns at a time: modified[i] = prod_j (v_j[i] +
 * delta_j*omega^i*beta + gamma) / prod_j (v_j[i] + beta*sigma_j[i] + gamma)
 * (denominators batch-inverted; zero inverts to zero), z[0] = last_z,
 * z[i] = z[i-1] * modified[i-1], the last `blinding` rows replaced by the
 * set's blinds, last_z = z[len - blinding - 1] (starts at one). */
void orc_permutation_products(orc_fr *out, const orc_fr *columns, const orc_fr *sigmas, size_t len,
                              size_t column_count, size_t chunk_len, const orc_fr *beta, const orc_fr *gamma,
                              const orc_fr *omega, const orc_fr *delta_powers, const orc_fr *blinds,
                              size_t blinding, int threads)
{
    const size_t sets = (column_count + chunk_len - 1) / chunk_len;
    orc_fr last_z; orc_fr_one(&last_z);
    orc_fr *num = (orc_fr *)malloc(len * sizeof(orc_fr));
    orc_fr *den = (orc_fr *)malloc(len * sizeof(orc_fr));
    orc_fr *pre = (orc_fr *)malloc(len * sizeof(orc_fr));
    if (threads <= 0) threads = 1;
    for (size_t s = 0; s < sets; s++) {
        size_t c0 = s * chunk_len, c1 = c0 + chunk_len < column_count ? c0 + chunk_len : column_count;
        const size_t nchunks = (size_t)threads * 4;
        const size_t chunk = (len + nchunks - 1) / nchunks;
        #pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
        for (size_t ch = 0; ch < nchunks; ch++) {
          size_t start = ch * chunk, end = start + chunk < len ? start + chunk : len;
          orc_fr wi;
          uint64_t e[4] = {(uint64_t)start, 0, 0, 0};
          orc_fr_pow(&wi, omega, e);
          for (size_t i = start; i < end; i++, orc_fr_mul(&wi, &wi, omega)) {
            orc_fr t, n1, d1;
            orc_fr_one(&n1);
            orc_fr_one(&d1);
            for (size_t j = c0; j < c1; j++) {
                const orc_fr *v = &columns[j * len + i];
                orc_fr_mul(&t, beta, &sigmas[j * len + i]);
                orc_fr_add(&t, &t, gamma);
                orc_fr_add(&t, &t, v);
                orc_fr_mul(&d1, &d1, &t);
                orc_fr_mul(&t, &delta_powers[j], &wi);
                orc_fr_mul(&t, &t, beta);
                orc_fr_add(&t, &t, gamma);
                orc_fr_add(&t, &t, v);
                orc_fr_mul(&n1, &n1, &t);
            }
            num[i] = n1;
            den[i] = d1;
          }
        }
        /* batch inversion of the denominators (ff BatchInvert): zeros stay zero */
        orc_fr acc; orc_fr_one(&acc);
        for (size_t i = 0; i < len; i++) {
            pre[i] = acc;
            if (!orc_fr_is_zero(&den[i])) orc_fr_mul(&acc, &acc, &den[i]);
        }
        orc_fr inv; orc_fr_inv(&inv, &acc);
        for (size_t i = len; i-- > 0;) {
            if (orc_fr_is_zero(&den[i])) continue;
            orc_fr t;
            orc_fr_mul(&t, &inv, &pre[i]);
            orc_fr_mul(&inv, &inv, &den[i]);
            den[i] = t;
        }
        orc_fr *z = out + s * len;
        z[0] = last_z;
        for (size_t i = 1; i < len; i++) {
            orc_fr m;
            orc_fr_mul(&m, &num[i - 1], &den[i - 1]);
            orc_fr_mul(&z[i], &z[i - 1], &m);
        }
        for (size_t i = len - blinding; i < len; i++) z[i] = blinds[s * blinding + (i - (len - blinding))];
        last_z = z[len - blinding - 1];
    }
    free(num); free(den); free(pre);
}

/* shifted_prefix_product (permutation/prover.rs:36-43): out[0] = initial,
 * out[i] = out[i-1] * values[i-1] */
void orc_shifted_prefix_product(orc_fr *out, const orc_fr *values, size_t len, const orc_fr *initial)
{
    if (len == 0) return;
    orc_fr acc = *initial, next;
    for (size_t i = 0; i < len; i++) {
        orc_fr_mul(&next, &acc, &values[i]);
        out[i] = acc;
        acc = next;
    }
}

/* Lookup product columns: vendor/midnight-proofs-0.8.1/src/plonk/lookup/
 * prover.rs:1075-1178 (construct_product).  Per lookup: denominators
 * (beta + a'_i)(gamma + s'_i), batch-inverted (zero stays zero), times
 * (a_i + beta)(s_i + gamma); z = 1, f_0, f_0 f_1, ... for the first
 * len - blinding rows, then the blinds.  Inputs lookup-major. */
void orc_lookup_products(orc_fr *out, const orc_fr *ci, const orc_fr *ct, const orc_fr *pi, const orc_fr *pt,
                         size_t len, size_t count, const orc_fr *beta, const orc_fr *gamma, const orc_fr *blinds,
                         size_t blinding)
{
    orc_fr *f = (orc_fr *)malloc(len * sizeof(orc_fr));
    for (size_t l = 0; l < count; l++) {
        const size_t o = l * len;
        for (size_t i = 0; i < len; i++) {
            orc_fr a, b;
            orc_fr_add(&a, beta, &pi[o + i]);
            orc_fr_add(&b, gamma, &pt[o + i]);
            orc_fr_mul(&f[i], &a, &b);
            orc_fr_inv(&f[i], &f[i]);                 /* 0 -> 0, as batch_invert */
            orc_fr_add(&a, &ci[o + i], beta);
            orc_fr_mul(&f[i], &f[i], &a);
            orc_fr_add(&b, &ct[o + i], gamma);
            orc_fr_mul(&f[i], &f[i], &b);
        }
        orc_fr acc;
        orc_fr_one(&acc);
        for (size_t i = 0; i + blinding < len; i++) {
            out[o + i] = acc;
            orc_fr_mul(&acc, &acc, &f[i]);
        }
        for (size_t t = 0; t < blinding; t++) out[o + len - blinding + t] = blinds[l * blinding + t];
    }
    free(f);
}

/* permute_expression_pair: vendor/midnight-proofs-0.8.1/src/plonk/lookup/
 * prover.rs:1573-1690, over the usable rows.  Inputs sorted (Fq's Ord, the
 * canonical integer order); each first occurrence copies its value to the
 * permuted table and consumes one equal table entry, walking the sorted
 * table with one pointer; the unconsumed table entries, ascending, fill the
 * repeated input rows taken from the last one back.  Returns 0 when an
 * input is not in the table. */
typedef struct {
    uint64_t c[4];
    orc_fr v;
} orc_keyed;

static int keyed_cmp(const void *pa, const void *pb)
{
    const orc_keyed *a = (const orc_keyed *)pa, *b = (const orc_keyed *)pb;
    for (int j = 3; j >= 0; j--) {
        if (a->c[j] < b->c[j]) return -1;
        if (a->c[j] > b->c[j]) return 1;
    }
    return 0;
}

static orc_keyed *keyed_sorted(const orc_fr *v, size_t n)
{
    orc_keyed *k = (orc_keyed *)malloc((n ? n : 1) * sizeof(orc_keyed));
    for (size_t i = 0; i < n; i++) {
        orc_fr_to_canonical(k[i].c, &v[i]);
        k[i].v = v[i];
    }
    qsort(k, n, sizeof(orc_keyed), keyed_cmp);   /* equal keys are equal values: stability is moot */
    return k;
}

void orc_fr_sort_canonical(orc_fr *out, const orc_fr *v, size_t n)
{
    orc_keyed *k = keyed_sorted(v, n);
    for (size_t i = 0; i < n; i++) out[i] = k[i].v;
    free(k);
}

int orc_lookup_permute(orc_fr *pin, orc_fr *ptab, const orc_fr *input, const orc_fr *table, size_t usable)
{
    orc_keyed *in = keyed_sorted(input, usable), *tb = keyed_sorted(table, usable);
    size_t *repeated = (size_t *)malloc((usable ? usable : 1) * sizeof(size_t));
    unsigned char *removed = (unsigned char *)calloc(usable ? usable : 1, 1);
    size_t nrep = 0, t = 0;
    int ok = 1;
    for (size_t r = 0; r < usable; r++) {
        pin[r] = in[r].v;
        if (r == 0 || keyed_cmp(&in[r], &in[r - 1]) != 0) {
            ptab[r] = in[r].v;
            while (t < usable && keyed_cmp(&tb[t], &in[r]) < 0) t++;
            if (t == usable || keyed_cmp(&tb[t], &in[r]) != 0) {
                ok = 0;
                break;
            }
            removed[t++] = 1;
        } else {
            repeated[nrep++] = r;
        }
    }
    for (size_t p = 0; ok && p < usable; p++) {
        if (removed[p]) continue;
        if (nrep == 0) {
            ok = 0;
            break;
        }
        ptab[repeated[--nrep]] = tb[p].v;
    }
    if (ok && nrep != 0) ok = 0;
    free(in);
    free(tb);
    free(repeated);
    free(removed);
    return ok;
}
