/*
 * CPU oracle for the BLS12-381 prover hot path (G1 MSM, Fr NTT / coset
 * transforms, quotient evaluation).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * backend is compared against, and the multithreaded CPU baseline that
 * bench.py times beside it.  Nothing in the shipped backend links, loads or
 * calls it.
 *
 * Value layouts follow the reference's ABI types
 * (midnight-sppark-msm/src/lib.rs:17-36, bls12_381_g1_msm.cu:336-349):
 *   Fr            4 x u64 little-endian limbs, Montgomery form, R = 2^256
 *   Fp            6 x u64 little-endian limbs, Montgomery form, R = 2^384
 *   G1 affine     x[6], y[6]; the point at infinity is all-zero (blst)
 *   G1 Jacobian   x[6], y[6], z[6]; x = X/Z^2, y = Y/Z^3, Z = 0 is infinity
 *
 * Parity pinning (what fixes each function independently of its own
 * formula) is documented per file and exercised by tests/test_oracle_*.py.
 */
#ifndef MIDNIGHT_ORACLE_H
#define MIDNIGHT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t l[4]; } orc_fr;
typedef struct { uint64_t l[6]; } orc_fp;
typedef struct { orc_fp x, y; } orc_g1a;
typedef struct { orc_fp x, y, z; } orc_g1j;

/* ---- Fr (scalar field) : vendor/midnight-curves-0.3.1/src/bls12_381/fq.rs ---- */
extern const orc_fr ORC_FR_MODULUS;      /* canonical, not Montgomery */
void orc_fr_add(orc_fr *r, const orc_fr *a, const orc_fr *b);
void orc_fr_sub(orc_fr *r, const orc_fr *a, const orc_fr *b);
void orc_fr_neg(orc_fr *r, const orc_fr *a);
void orc_fr_mul(orc_fr *r, const orc_fr *a, const orc_fr *b);
void orc_fr_sqr(orc_fr *r, const orc_fr *a);
void orc_fr_pow(orc_fr *r, const orc_fr *a, const uint64_t e[4]);
void orc_fr_inv(orc_fr *r, const orc_fr *a);           /* 0 -> 0 */
void orc_fr_one(orc_fr *r);
void orc_fr_from_u64(orc_fr *r, uint64_t v);
void orc_fr_to_canonical(uint64_t out[4], const orc_fr *a);
int  orc_fr_from_canonical(orc_fr *r, const uint64_t in[4]);   /* 1 if < r */
int  orc_fr_eq(const orc_fr *a, const orc_fr *b);
int  orc_fr_is_zero(const orc_fr *a);
/* Fq::from_uniform_bytes (fq.rs:704-726): a0*R2 + a1*R3 */
void orc_fr_from_uniform_bytes(orc_fr *r, const uint8_t bytes[64]);
void orc_fr_inner_product(orc_fr *r, const orc_fr *a, const orc_fr *b, size_t n);
void orc_fr_mul_add_many(orc_fr *out, const orc_fr *a, const orc_fr *b, const orc_fr *c, size_t n);
void orc_fr_lincomb(orc_fr *out, const orc_fr *polys, const orc_fr *s, size_t m, size_t n);
/* named constants, Montgomery form (fq.rs:60-135, 476-505, 693-702) */
void orc_fr_root_of_unity(orc_fr *r);      /* 2^32-th root */
void orc_fr_zeta(orc_fr *r);               /* cube root of unity ZETA */
void orc_fr_delta(orc_fr *r);

/* ---- Fp (base field) : vendor/midnight-curves-0.3.1/src/bls12_381/fp.rs ---- */
void orc_fp_add(orc_fp *r, const orc_fp *a, const orc_fp *b);
void orc_fp_sub(orc_fp *r, const orc_fp *a, const orc_fp *b);
void orc_fp_neg(orc_fp *r, const orc_fp *a);
void orc_fp_mul(orc_fp *r, const orc_fp *a, const orc_fp *b);
void orc_fp_sqr(orc_fp *r, const orc_fp *a);
void orc_fp_inv(orc_fp *r, const orc_fp *a);
void orc_fp_inv_vartime(orc_fp *r, const orc_fp *a);
void orc_fp_one(orc_fp *r);
void orc_fp_to_canonical(uint64_t out[6], const orc_fp *a);
int  orc_fp_from_canonical(orc_fp *r, const uint64_t in[6]);
int  orc_fp_eq(const orc_fp *a, const orc_fp *b);
int  orc_fp_is_zero(const orc_fp *a);
int  orc_fp_sqrt(orc_fp *r, const orc_fp *a);   /* 1 if a is a square */

/* ---- G1 : vendor/midnight-curves-0.3.1/src/bls12_381/g1.rs ---- */
void orc_g1_generator(orc_g1a *r);
void orc_g1j_identity(orc_g1j *r);
int  orc_g1j_is_identity(const orc_g1j *p);
int  orc_g1a_is_identity(const orc_g1a *p);
void orc_g1j_from_affine(orc_g1j *r, const orc_g1a *p);
void orc_g1j_to_affine(orc_g1a *r, const orc_g1j *p);
void orc_g1j_dbl(orc_g1j *r, const orc_g1j *p);
void orc_g1j_add(orc_g1j *r, const orc_g1j *p, const orc_g1j *q);
void orc_g1j_add_affine(orc_g1j *r, const orc_g1j *p, const orc_g1a *q);
void orc_g1j_neg(orc_g1j *r, const orc_g1j *p);
int  orc_g1j_eq(const orc_g1j *p, const orc_g1j *q);
int  orc_g1a_is_on_curve(const orc_g1a *p);
/* scalar is an Fr in Montgomery form; result is [s]P */
void orc_g1_mul(orc_g1j *r, const orc_g1a *p, const orc_fr *s);
void orc_g1j_mul(orc_g1j *r, const orc_g1j *p, const orc_fr *s);
/* batch normalisation Jacobian -> affine (one inversion) */
void orc_g1_batch_to_affine(orc_g1a *out, const orc_g1j *in, size_t n);
/* ZCash serialisation (48-byte compressed, 96-byte uncompressed) */
void orc_g1_compress(uint8_t out[48], const orc_g1a *p);
int  orc_g1_decompress(orc_g1a *r, const uint8_t in[48]);
size_t orc_g1_decompress_many(orc_g1a *out, const uint8_t *in, size_t n, int threads);
int  orc_g1_from_uncompressed(orc_g1a *r, const uint8_t in[96]);
void orc_g1_to_uncompressed(uint8_t out[96], const orc_g1a *p);

/* ---- MSM ---- */
/* straightforward sum of [s_i]P_i by double-and-add; small n only */
void orc_msm_naive(orc_g1j *out, const orc_g1a *pts, const orc_fr *sc, size_t n);
/* Pippenger with Booth-signed windows, windows spread over threads
 * (restates midnight-curves msm.rs:477 msm_best; c = ceil(ln n)) */
void orc_msm_pippenger(orc_g1j *out, const orc_g1a *pts, const orc_fr *sc,
                       size_t n, int threads);
void orc_msm_best(orc_g1j *out, const orc_g1a *pts, const orc_fr *sc, size_t n, int threads);
/* Booth digit of window w (msm.rs:13 get_booth_index); canonical LE bytes */
int32_t orc_booth_index(size_t window_index, size_t window_size, const uint8_t el[32]);
/* SRS helper: out[i] = [tau^i]G for i < n (threaded) */
void orc_srs_powers(orc_g1a *out, const orc_fr *tau, size_t n, int threads);
/* out[i] = [t_i]G (threaded) */
void orc_g1_fixed_base_batch(orc_g1a *out, const orc_fr *t, size_t n, int threads);

/* ---- NTT / domain ---- */
/* best_fft restated (midnight-curves fft.rs:30-112,180-216): in-place DIT,
 * natural order in and out, no scaling */
void orc_best_fft(orc_fr *a, const orc_fr *omega, uint32_t log_n, int threads);
/* omega for 2^k: ROOT_OF_UNITY^(2^(32-k)) (domain.rs:71-86) */
void orc_omega(orc_fr *r, uint32_t log_n);
/* iNTT with 1/n scaling (domain.rs:579 ifft) */
void orc_ifft(orc_fr *a, uint32_t log_n, int threads);
/* forward NTT with canonical omega (coeff_to_lagrange) */
void orc_fft(orc_fr *a, uint32_t log_n, int threads);
/* a[i] *= {1, first, second}[i % 3] (domain.rs:544-577 distribute_powers_zeta) */
void orc_mul_zeta(orc_fr *a, size_t n, const orc_fr *first, const orc_fr *second);
/* a[i] *= f[i % m] (domain.rs:476 divide_by_vanishing_poly) */
void orc_mul_repeating(orc_fr *a, size_t n, const orc_fr *f, size_t m);
void orc_mul_scalar(orc_fr *a, size_t n, const orc_fr *s);
/* EvaluationDomain::new(j, k) (domain.rs:54-159).  t_evals gets
 * 2^(extended_k - k) entries (already inverted). */
typedef struct {
    uint32_t k, extended_k, j;
    orc_fr omega, omega_inv, extended_omega, extended_omega_inv;
    orc_fr ifft_divisor, extended_ifft_divisor;
    orc_fr g_coset, g_coset_inv, barycentric_weight;
} orc_domain;
void orc_domain_new(orc_domain *d, orc_fr *t_evals, uint32_t j, uint32_t k);
/* coeff_to_extended (domain.rs:325): zeta-distribute, zero pad, forward NTT */
void orc_coeff_to_extended(orc_fr *out, const orc_fr *coeffs, const orc_domain *d, int threads);
/* extended_to_coeff (domain.rs:410): iNTT(ext) * 1/N_ext, zeta^-1 distribute */
void orc_extended_to_coeff(orc_fr *a, const orc_domain *d, int threads);

/* ---- Quotient numerator (plonk/evaluation.rs) ---- */
/* Compiled expression graph, same layout as the GPU ABI's RawGraph
 * (bls12_381_g1_msm.cu:351-376).  Source kinds: 0 constant, 1 intermediate,
 * 2 fixed, 3 advice, 4 instance (a = column, b = rotation slot), 5 challenge,
 * 6 beta, 7 gamma, 8 theta, 9 trash challenge, 10 y, 11 previous value.
 * Opcodes: 0 add, 1 sub, 2 mul, 3 square, 4 double, 5 negate, 6 Horner
 * (value = lhs; value = value*rhs + part for each part), 7 store. */
typedef struct { uint32_t kind, a, b; } orc_value_source;
typedef struct {
    uint32_t op, target, parts_offset, parts_len;
    orc_value_source lhs, rhs;
} orc_calculation;
typedef struct {
    const orc_fr *constants; size_t constants_len;
    const int32_t *rotations; size_t rotations_len;
    const orc_calculation *calculations; size_t calculations_len;
    const orc_value_source *horner_parts; size_t horner_parts_len;
    uint32_t num_intermediates;
} orc_graph;
typedef struct { uint32_t kind, index; } orc_column_ref;   /* 0 fixed, 1 advice, 2 instance */

/* Everything one numerator evaluation reads.  Column-major polynomial
 * arrays: column c occupies [c*len, (c+1)*len). */
typedef struct {
    size_t len;
    int run_custom;                /* 0: start from initial_out, skip gates */
    const orc_fr *initial_out;
    const orc_fr *fixed; size_t fixed_cols;
    const orc_fr *advice; size_t advice_cols;
    const orc_fr *instance; size_t instance_cols;
    const orc_fr *challenges; size_t challenge_len;
    orc_fr beta, gamma, theta, trash_challenge, y;
    orc_fr omega;                  /* generator of the evaluated domain */
    orc_fr g_coset;                /* delta_start = beta * g_coset */
    int32_t rot_scale, last_rotation;
    uint32_t chunk_len;
    const orc_fr *l0, *l_last, *l_active;
    const orc_fr *permutation_products; size_t permutation_sets;
    const orc_column_ref *permutation_columns; size_t permutation_columns_len;
    const orc_fr *permutation_pk_cosets;
    const orc_fr *lookup_product_cosets, *lookup_input_cosets, *lookup_table_cosets;
    size_t lookup_count;
    const orc_fr *trash_cosets;
    const uint32_t *trash_selector_fixed_indices; size_t trash_count;
    orc_graph custom_graph;
    const orc_graph *lookup_graphs;
    const orc_graph *trash_graphs;
} orc_numerator_inputs;

/* GraphEvaluator::evaluate over rows 0..len (evaluation.rs:3797-3845);
 * previous[i] is the PreviousValue of row i (NULL: zero). */
void orc_graph_eval_rows(orc_fr *out, const orc_graph *g, const orc_numerator_inputs *in,
                         const orc_fr *previous, int threads);
/* Evaluator::evaluate_numerator for one proof (evaluation.rs:3114-3528) */
void orc_eval_numerator(orc_fr *out, const orc_numerator_inputs *in, int threads);

/* Permutation grand products (permutation/prover.rs:355-515); out gets
 * ceil(column_count / chunk_len) sets of len values. */
void orc_permutation_products(orc_fr *out, const orc_fr *columns, const orc_fr *sigmas, size_t len,
                              size_t column_count, size_t chunk_len, const orc_fr *beta, const orc_fr *gamma,
                              const orc_fr *omega, const orc_fr *delta_powers, const orc_fr *blinds,
                              size_t blinding, int threads);

/* Lookup product columns (lookup/prover.rs:1075-1178); inputs and out lookup-major, count * len values;
 * blinds: count * blinding values. */
void orc_lookup_products(orc_fr *out, const orc_fr *ci, const orc_fr *ct, const orc_fr *pi, const orc_fr *pt,
                         size_t len, size_t count, const orc_fr *beta, const orc_fr *gamma, const orc_fr *blinds,
                         size_t blinding);

/* values sorted by canonical value (Fq's Ord) */
void orc_fr_sort_canonical(orc_fr *out, const orc_fr *v, size_t n);
/* permute_expression_pair (lookup/prover.rs:1573-1690) over `usable` rows; 0 if an input is not in the table */
int orc_lookup_permute(orc_fr *pin, orc_fr *ptab, const orc_fr *input, const orc_fr *table, size_t usable);

/* ---- Poseidon (bls12_381_g1_msm.cu:2018-2216): width 3, 8 full + 60 partial rounds, x^5 ---- */
/* permutation in place; w (128 values) or NULL */
void orc_poseidon_permutation(orc_fr state[3], const orc_fr *mds, const orc_fr *rc, orc_fr *w);
/* dependent chain of len two-permutation hashes; out (len) and witness (len * 256) may be NULL */
void orc_poseidon_hash_chain(orc_fr *out, orc_fr *witness, size_t len, const orc_fr *initial,
                             const orc_fr *tail0, const orc_fr *tail1, const orc_fr *input_len,
                             const orc_fr *mds, const orc_fr *rc);
/* 8 * domain_len advice values from the chain witness, metadata (11) and packed layout (128) */
void orc_poseidon_advice_from_witness(orc_fr *out, const orc_fr *witness, size_t hash_count, size_t domain_len,
                                      const orc_fr *metadata, const uint16_t *layout);

/* shifted_prefix_product (permutation/prover.rs:36-43); out may not alias values */
void orc_shifted_prefix_product(orc_fr *out, const orc_fr *values, size_t len, const orc_fr *initial);

/* ---- coefficient-form polynomials (utils/arithmetic.rs) ---- */
void orc_poly_eval(orc_fr *out, const orc_fr *poly, size_t n, const orc_fr *point);
/* q[0 .. n-2] = a / (X - b) without remainder, q[n-1] = 0; q may alias a */
void orc_kate_division(orc_fr *q, const orc_fr *a, size_t n, const orc_fr *b);

/* ---- RNG : rand_core 0.6 seed_from_u64 (PCG32) + rand_chacha 0.3 ChaCha20 ---- */
typedef struct {
    uint32_t key[8];
    uint64_t counter;       /* next block counter */
    uint32_t buf[64];       /* four blocks, as rand_chacha buffers them */
    uint32_t idx;           /* next word in buf; 64 = empty */
} orc_rng;
void orc_rng_seed_from_u64(orc_rng *r, uint64_t seed);
void orc_rng_from_seed(orc_rng *r, const uint8_t seed[32]);
uint32_t orc_rng_next_u32(orc_rng *r);
uint64_t orc_pcg32_advance(uint64_t state, uint64_t inc);
uint32_t orc_pcg32_output(uint64_t state);
uint64_t orc_rng_next_u64(orc_rng *r);
/* Fq::random (fq.rs:373-387): rejection sampling of 255-bit values */
void orc_fr_random(orc_fr *r, orc_rng *rng);
void orc_fr_random_many(orc_fr *out, size_t n, orc_rng *rng);
/* ChaCha20 block function (RFC 7539 2.3 with a 64-bit counter, zero nonce) */
void orc_chacha20_block(uint32_t out[16], const uint32_t key[8], uint64_t counter, const uint32_t nonce[2]);

#ifdef __cplusplus
}
#endif
#endif
