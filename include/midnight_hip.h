/*
 * midnight_hip.h -- C ABI of the MI355X (gfx950) backend for the Midnight
 * prover hot path: BLS12-381 G1 MSM, Fr NTT / coset transforms, and the
 * quotient numerator on the extended domain.
 *
 * Each `midnight_sppark_*` entry point keeps the name, argument order,
 * value layout and error convention of the reference CUDA backend
 * (midnight-sppark-msm/src/bls12_381_g1_msm.cu) so that the Rust FFI block in
 * midnight-sppark-msm/src/lib.rs:763-1362 links against this library
 * unchanged.  Entry points prefixed `midnight_hip_` have no counterpart in
 * the reference; each says who it is for.
 *
 * Conventions (bls12_381_g1_msm.cu:336-349, lib.rs:17-36):
 *   RawFr            4 x u64 LE limbs, Montgomery form, R = 2^256
 *   RawG1Affine      x[6], y[6] Montgomery (R = 2^384); identity = all zero
 *   RawG1Projective  Jacobian x[6], y[6], z[6]; identity has z = 0.  Callers
 *                    compare results after normalisation.
 *   RustError        returned by value; code 0 = success, otherwise a
 *                    hipError_t value or -1; message is malloc'd (the Rust
 *                    side frees it with libc free), NULL on success
 *                    (vendor/sppark-0.1.15/sppark/util/rusterror.h).
 *   Host pointers are caller-owned and used synchronously.  `void*` handles
 *   (MSM contexts, device vectors, numerator contexts) are freed with the
 *   matching *_free function.
 *   Lanes: the lane is thread-local (bls12_381_g1_msm.cu:48); all work of
 *   one lane is ordered on that lane's stream.
 */
#ifndef MIDNIGHT_HIP_H
#define MIDNIGHT_HIP_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct RustError {
    int code;
    char *message;
} RustError;

typedef struct RawFr {
    uint64_t limbs[4];
} RawFr;

typedef struct RawG1Affine {
    uint64_t x[6];
    uint64_t y[6];
} RawG1Affine;

typedef struct RawG1Projective {
    uint64_t x[6];
    uint64_t y[6];
    uint64_t z[6];
} RawG1Projective;

/* bls12_381_g1_msm.cu:50-61.  window_bits 0 = automatic; the other knobs
 * are validated as the reference validates them (:85-110) and mapped onto
 * this backend's own tuning (window, engines). */
typedef struct RawMsmConfig {
    uint32_t domain_k;
    uint32_t min_msm_scalars;
    uint32_t unfiltered_min_scalars;
    uint32_t window_bits;
    uint32_t sort_grid_divisor;
    uint32_t sort_window_lanes;
    uint32_t integration_blocks;
    uint32_t batch_lanes;
    uint32_t reuse_event;
    uint32_t device_reduce;
} RawMsmConfig;

/* bls12_381_g1_msm.cu:63-68 */
typedef struct RawMsmSlice {
    size_t scalar_offset;
    size_t point_offset;
    size_t len;
    size_t nonzero_len;
} RawMsmSlice;

/* ---------------- lanes ---------------- */

/* replaces bls12_381_g1_msm.cu:134 midnight_sppark_execution_lane_set */
RustError midnight_sppark_execution_lane_set(uint32_t lane);

/* ---------------- G1 MSM ---------------- */

/* replaces bls12_381_g1_msm.cu:6164 */
RustError midnight_sppark_bls12_381_g1_msm(RawG1Projective *out, const RawG1Affine *points,
                                           const RawFr *scalars, size_t len);
/* replaces bls12_381_g1_msm.cu:6202 */
RustError midnight_sppark_bls12_381_g1_msm_batch(RawG1Projective *out, const RawG1Affine *points,
                                                 const RawFr *scalars, size_t npoints, size_t nmsms);
/* replaces bls12_381_g1_msm.cu:6247 */
RustError midnight_sppark_bls12_381_g1_msm_context_create_with_config(void **out,
                                                                      const RawG1Affine *points,
                                                                      size_t npoints,
                                                                      RawMsmConfig config);
/* replaces bls12_381_g1_msm.cu:6271 */
RustError midnight_sppark_bls12_381_g1_msm_context_create(void **out, const RawG1Affine *points,
                                                          size_t npoints);
/* replaces bls12_381_g1_msm.cu:6283 */
void midnight_sppark_bls12_381_g1_msm_context_free(void *context);
/* replaces bls12_381_g1_msm.cu:6288 */
RustError midnight_sppark_bls12_381_g1_msm_context_batch(RawG1Projective *out, void *context,
                                                         const RawFr *scalars, size_t npoints,
                                                         size_t nmsms);
/* replaces bls12_381_g1_msm.cu:6310 */
RustError midnight_sppark_bls12_381_g1_msm_context_device_batch(RawG1Projective *out,
                                                                void *context,
                                                                const void *device_vec,
                                                                size_t scalar_offset,
                                                                size_t npoints, size_t nmsms);
/* replaces bls12_381_g1_msm.cu:6346 */
RustError midnight_sppark_bls12_381_g1_msm_context_device_ragged(RawG1Projective *out,
                                                                 void *context,
                                                                 const void *device_vec,
                                                                 const RawMsmSlice *slices,
                                                                 size_t count);
/* replaces bls12_381_g1_msm.cu:6426 */
RustError midnight_sppark_bls12_381_g1_msm_context_device_scrambled_batch(
    RawG1Projective *out, void *context, const void *device_vec, size_t scalar_offset,
    size_t npoints, size_t nmsms, const void *mask_device_vec);

/* ---------------- Fr NTT / coset ---------------- */

/* Test hooks (bls12_381_g1_msm.cu:6379-6424, the reference's "test-hooks"
 * feature; Rust tests lib.rs:5088-5209): hold a device-vector producer back
 * by a host callback that sleeps on the lane's stream.  mask_delay applies
 * to the calling thread's next scrambled batches (0 turns it off);
 * delayed_scale multiplies a device vector by `scale` behind the delay and
 * returns before it has run. */
void midnight_sppark_test_mask_delay(uint32_t delay_ms);
RustError midnight_sppark_test_delayed_scale(void *raw_device, RawFr scale, uint32_t delay_ms);
/* Test hook (no counterpart in the reference): caps the bucket-tree levels
 * of the calling thread's MSMs that read their fullest bucket back (the
 * context and one-shot entry points); the last reduction pass then sums
 * all the partials deeper buckets still hold, which tests that pass on
 * buckets of any depth.  0xffffffff lifts the cap. */
void midnight_hip_test_msm_tree_levels(uint32_t max_levels);
/* replaces bls12_381_g1_msm.cu:6488.  In-place natural-order NTT of 2^log_n
 * values; forward uses omega = ROOT_OF_UNITY^(2^(32-log_n)), inverse uses
 * omega^-1 and includes the 1/n factor. */
RustError midnight_sppark_bls12_381_fr_ntt(RawFr *values, uint32_t log_n, bool inverse);
/* replaces bls12_381_g1_msm.cu:3447: NTT then v[i] *= scale * {1,first,second}[i % 3] */
RustError midnight_sppark_bls12_381_fr_ntt_zeta(RawFr *values, uint32_t log_n, bool inverse,
                                                RawFr scale, RawFr first, RawFr second);
/* replaces bls12_381_g1_msm.cu:3497: v[i] *= factors[i % factor_len], inverse NTT,
 * then the ntt_zeta post-processing */
RustError midnight_sppark_bls12_381_fr_divide_ntt_zeta(RawFr *values, uint32_t log_n,
                                                       const RawFr *factors, size_t factor_len,
                                                       RawFr scale, RawFr first, RawFr second);
/* replaces bls12_381_g1_msm.cu:2542: per poly, a[i] *= {1,first,second}[i % 3],
 * zero-pad to 2^extended_log_n, forward NTT */
RustError midnight_sppark_bls12_381_fr_coeff_to_extended_batch(RawFr *out, const RawFr *inputs,
                                                               size_t input_len, size_t count,
                                                               uint32_t extended_log_n,
                                                               RawFr first, RawFr second);
/* replaces bls12_381_g1_msm.cu:2604 (result stays on device) */
RustError midnight_sppark_bls12_381_fr_device_coeff_to_extended_batch(
    void **out, const RawFr *inputs, size_t input_len, size_t count, uint32_t extended_log_n,
    RawFr first, RawFr second);
/* replaces bls12_381_g1_msm.cu:2675 */
RustError midnight_sppark_bls12_381_fr_device_coeff_to_extended_from_device_batch(
    void **out, const void *source, size_t input_len, size_t count, uint32_t extended_log_n,
    RawFr first, RawFr second);
/* replaces bls12_381_g1_msm.cu:3281 */
RustError midnight_sppark_bls12_381_fr_device_vec_ntt_zeta(void *device_vec, uint32_t log_n,
                                                           bool inverse, RawFr scale, RawFr first,
                                                           RawFr second);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation): midnight_sppark_bls12_381_fr_device_vec_ntt_zeta with
 * the same result, its four-step staging taken from a static 32 MiB scratch
 * array per GPU instead of an allocation.  log_n at most 20 (larger is an
 * error).  Calls that overlap on one GPU share that array and race. */
RustError midnight_hip_fr_device_vec_ntt_zeta_noalloc(void *device_vec, uint32_t log_n, bool inverse,
                                                      RawFr scale, RawFr first, RawFr second);
/* replaces bls12_381_g1_msm.cu:3331 */
RustError midnight_sppark_bls12_381_fr_device_vec_ntt_zeta_batch(void *device_vec,
                                                                 uint32_t log_n, size_t count,
                                                                 bool inverse, RawFr scale,
                                                                 RawFr first, RawFr second);
/* replaces bls12_381_g1_msm.cu:3385 */
RustError midnight_sppark_bls12_381_fr_device_vec_divide_ntt_zeta(void *device_vec,
                                                                  uint32_t log_n,
                                                                  const RawFr *factors,
                                                                  size_t factor_len, RawFr scale,
                                                                  RawFr first, RawFr second);

/* ---------------- Fr vector ops ---------------- */

/* replaces bls12_381_g1_msm.cu:6504 */
RustError midnight_sppark_bls12_381_fr_mul_assign(RawFr *values, RawFr scalar, size_t len);
/* replaces bls12_381_g1_msm.cu:6521 */
RustError midnight_sppark_bls12_381_fr_mul_assign_zeta(RawFr *values, RawFr first, RawFr second,
                                                       size_t len);
/* replaces bls12_381_g1_msm.cu:6544 */
RustError midnight_sppark_bls12_381_fr_mul_assign_repeating(RawFr *values, const RawFr *factors,
                                                            size_t factor_len, size_t len);

/* ---------------- device vectors ---------------- */

/* replaces bls12_381_g1_msm.cu:2741 */
RustError midnight_sppark_bls12_381_fr_device_vec_upload(void **out, const RawFr *values,
                                                         size_t len);
/* replaces bls12_381_g1_msm.cu:2774 */
RustError midnight_sppark_bls12_381_fr_device_vec_upload_many(void **out,
                                                              const RawFr *const *values,
                                                              const size_t *lengths, size_t count);
/* replaces bls12_381_g1_msm.cu:3074 */
RustError midnight_sppark_bls12_381_fr_device_vec_download(RawFr *out, const void *device_vec,
                                                           size_t len);
/* replaces bls12_381_g1_msm.cu:3442 */
void midnight_sppark_bls12_381_fr_device_vec_free(void *device_vec);
/* replaces bls12_381_g1_msm.cu:3181: split the quotient into piece_count
 * limbs of poly_len - 1 coefficients padded to poly_len, then add the
 * blinding telescope (piece i gets +blind_i at index poly_len - 1 and piece
 * i + 1 gets -blind_i at index 0) */
RustError midnight_sppark_bls12_381_fr_device_quotient_pieces(void **out,
                                                              const void *quotient_vec,
                                                              size_t poly_len, size_t piece_count,
                                                              const RawFr *blinds,
                                                              size_t blind_count);

/* replaces bls12_381_g1_msm.cu:2878: device_vec[indices[i]] = values[i] */
RustError midnight_sppark_bls12_381_fr_device_vec_scatter_host(void *device_vec, const uint32_t *indices,
                                                               const RawFr *values, size_t count);
/* replaces bls12_381_g1_msm.cu:2916: new vector = prefix ++ suffix (host) */
RustError midnight_sppark_bls12_381_fr_device_vec_concat_host(void **out, const void *prefix_device_vec,
                                                              const RawFr *suffix, size_t suffix_len);
/* replaces bls12_381_g1_msm.cu:2964: new vector = sources in order ++ suffix (host) */
RustError midnight_sppark_bls12_381_fr_device_vec_concat_many_host(void **out,
                                                                   const void *const *source_device_vecs,
                                                                   size_t source_count, const RawFr *suffix,
                                                                   size_t suffix_len);
/* replaces bls12_381_g1_msm.cu:3025: new vector = polynomials indices[i] (poly_len each) of the source */
RustError midnight_sppark_bls12_381_fr_device_vec_gather_polys(void **out, const void *source_device_vec,
                                                               size_t poly_len, const uint32_t *indices,
                                                               size_t index_count);
/* replaces bls12_381_g1_msm.cu:3101: per polynomial of [scalar_offset, +poly_len*count), starts[i] = first
 * nonzero index (poly_len if none) and nonzero_counts[i] (optional); out receives the suffixes
 * poly[starts[i]..] back to back (out == NULL and out_capacity == 0: only *out_len) */
RustError midnight_sppark_bls12_381_fr_device_vec_download_nonzero_suffixes(
    RawFr *out, size_t out_capacity, size_t *out_len, size_t *starts, size_t *nonzero_counts,
    const void *device_vec, size_t scalar_offset, size_t poly_len, size_t count);

/* ---------------- polynomial combinations and multi-open ---------------- */

/* (source, index): source 0 dynamic, 1 fixed, 2 extra (bls12_381_g1_msm.cu:383-386) */
typedef struct RawDevicePolyRef {
    uint32_t source;
    uint32_t index;
} RawDevicePolyRef;

/* replaces bls12_381_g1_msm.cu:3558: out = sum_j scalars[j] * polys[j] (host arrays) */
RustError midnight_sppark_bls12_381_fr_poly_lincomb(RawFr *out, const RawFr *polys, size_t poly_len,
                                                    size_t poly_count, const RawFr *scalars);
/* replaces bls12_381_g1_msm.cu:3234: new device vector sum_j scalars[j] * polys[j] */
RustError midnight_sppark_bls12_381_fr_device_poly_lincomb(void **out, const void *polys_vec, size_t poly_len,
                                                           size_t poly_count, const RawFr *scalars,
                                                           size_t scalar_count);
/* replaces bls12_381_g1_msm.cu:3985: out[j] = poly_j(point), contiguous polynomials */
RustError midnight_sppark_bls12_381_fr_device_eval_polys(RawFr *out, const void *device_vec, size_t poly_len,
                                                         size_t poly_count, RawFr point);
/* replaces bls12_381_g1_msm.cu:3600: q_out[k] = sum over items t of q_offsets[k]..q_offsets[k+1] of
 * q_coefficients[t] * polys[q_poly_indices[t]]; eval_out[i] = polys[query_poly_indices[i]] at
 * points[query_point_indices[i]] (poly/kzg/mod.rs:1978-2570) */
RustError midnight_sppark_bls12_381_fr_multiopen_prepare(
    RawFr *q_out, RawFr *eval_out, const RawFr *polys, size_t poly_len, size_t poly_count,
    const uint32_t *q_offsets, size_t q_offsets_len, const uint32_t *q_poly_indices, const RawFr *q_coefficients,
    size_t q_items_len, const RawFr *points, size_t point_count, const uint32_t *query_poly_indices,
    const uint32_t *query_point_indices, size_t query_count);
/* replaces bls12_381_g1_msm.cu:3723: the same over resident polynomials (poly_refs name a source vector
 * and a polynomial index); q polynomials to q_out (host) and/or *q_device_out */
RustError midnight_sppark_bls12_381_fr_multiopen_prepare_device(
    RawFr *q_out, void **q_device_out, RawFr *eval_out, const void *dynamic_source, const void *fixed_source,
    const void *extra_source, size_t poly_len, const RawDevicePolyRef *poly_refs, size_t poly_count,
    const uint32_t *q_offsets, size_t q_offsets_len, const uint32_t *q_poly_indices, const RawFr *q_coefficients,
    size_t q_items_len, const RawFr *points, size_t point_count, const uint32_t *query_poly_indices,
    const uint32_t *query_point_indices, size_t query_count);
/* replaces bls12_381_g1_msm.cu:3873: f = sum_k s_k * (q_k divided by (X - a) for each point a of
 * points[point_offsets[k]..point_offsets[k+1]], remainders dropped) */
RustError midnight_sppark_bls12_381_fr_multiopen_f_poly_device(
    void **out, const void *q_device_vec, size_t poly_len, const uint32_t *point_offsets,
    size_t point_offsets_len, const RawFr *points, size_t point_count, const RawFr *combination_scalars,
    size_t scalar_count);
/* replaces bls12_381_g1_msm.cu:4025: pi = (final - final(point)) / (X - point) with
 * final = sum_k s_k q_k + s_last f; poly_len - 1 coefficients */
RustError midnight_sppark_bls12_381_fr_multiopen_pi_device(void **out, const void *q_device_vec,
                                                           const void *f_device_vec, size_t poly_len,
                                                           const RawFr *combination_scalars, size_t scalar_count,
                                                           RawFr point);
/* replaces bls12_381_g1_msm.cu:6582: values[i] = 1 / values[i] (0 stays 0) */
RustError midnight_sppark_bls12_381_fr_invert_assign(RawFr *values, size_t len);

/* ---------------- quotient numerator ---------------- */

/* Compiled expression graph (bls12_381_g1_msm.cu:351-376; built by the
 * prover's graph compiler, vendor/midnight-proofs-0.8.1/src/plonk/
 * evaluation.rs:677-800).  Source kinds: 0 constant[a], 1 slot[a],
 * 2 fixed / 3 advice / 4 instance column a at rotations[b], 5 challenge[a],
 * 6 beta, 7 gamma, 8 theta, 9 trash challenge, 10 y, 11 previous value
 * (zero in every evaluation below).  Opcodes: 0 add, 1 sub, 2 mul,
 * 3 square, 4 double, 5 negate, 6 Horner (value = lhs, then
 * value = value * rhs + part over horner_parts[parts_offset ..
 * parts_offset + parts_len]), 7 store.  The result of a graph is the value
 * of its last calculation.  num_intermediates (<= 256) is the slot count;
 * a calculation whose target is not below it is rejected.  Out-of-range
 * constant, column, rotation and challenge indices read as zero. */
typedef struct RawValueSource {
    uint32_t kind;
    uint32_t a;
    uint32_t b;
} RawValueSource;

typedef struct RawCalculation {
    uint32_t op;
    uint32_t target;
    uint32_t parts_offset;
    uint32_t parts_len;
    RawValueSource lhs;
    RawValueSource rhs;
} RawCalculation;

typedef struct RawGraph {
    const RawFr *constants;
    size_t constants_len;
    const int32_t *rotations;
    size_t rotations_len;
    const RawCalculation *calculations;
    size_t calculations_len;
    const RawValueSource *horner_parts;
    size_t horner_parts_len;
    uint32_t num_intermediates;
} RawGraph;

/* kind 0 fixed, 1 advice, 2 instance (bls12_381_g1_msm.cu:378-381) */
typedef struct RawColumnRef {
    uint32_t kind;
    uint32_t index;
} RawColumnRef;

/* replaces bls12_381_g1_msm.cu:5196 midnight_sppark_bls12_381_eval_numerator.
 * Host buffers, column-major (column c at [c*len, (c+1)*len)).  Per row idx
 * of the evaluated domain (generator omega, rotations scaled by rot_scale,
 * indices taken modulo len) the value is folded as value = value*y + term:
 * custom graph (or initial_out[idx] when run_custom == 0), then the
 * permutation terms of permutation_sets product columns (chunk_len columns
 * per set, delta_start * omega^idx for the first identity column, times
 * delta_const per column), then each lookup's five terms, then each trash
 * term -- the fold order of evaluation.rs:3114-3528. */
RustError midnight_sppark_bls12_381_eval_numerator(
    RawFr *out, size_t len, const RawFr *initial_out, uint32_t run_custom,
    const RawFr *fixed, size_t fixed_cols, const RawFr *advice, size_t advice_cols,
    const RawFr *instance, size_t instance_cols, const RawFr *challenges, size_t challenge_len,
    RawFr beta, RawFr gamma, RawFr theta, RawFr trash_challenge, RawFr y, RawFr omega,
    RawFr delta_start, RawFr delta_const, int32_t rot_scale, int32_t last_rotation,
    uint32_t chunk_len, const RawFr *l0, const RawFr *l_last, const RawFr *l_active,
    const RawFr *permutation_products, size_t permutation_sets,
    const RawColumnRef *permutation_columns, size_t permutation_columns_len,
    const RawFr *permutation_pk_cosets, size_t permutation_pk_cosets_len,
    const RawFr *lookup_product_cosets, const RawFr *lookup_input_cosets,
    const RawFr *lookup_table_cosets, size_t lookup_count, const RawFr *trash_cosets,
    const uint32_t *trash_selector_fixed_indices, size_t trash_count, RawGraph custom_graph,
    const RawGraph *lookup_graphs, const RawGraph *trash_graphs);

/* replaces bls12_381_g1_msm.cu:5434: out[g*len + idx] = graph g at row idx */
RustError midnight_sppark_bls12_381_eval_graph_batch(
    RawFr *out, size_t len, const RawFr *fixed, size_t fixed_cols, const RawFr *advice,
    size_t advice_cols, const RawFr *instance, size_t instance_cols, const RawFr *challenges,
    size_t challenge_len, RawFr beta, RawFr gamma, RawFr theta, RawFr trash_challenge, RawFr y,
    int32_t rot_scale, const RawGraph *graphs, size_t graph_count);

/* replaces bls12_381_g1_msm.cu:5517: keeps the key-dependent inputs of the
 * numerator (fixed and sigma cosets, Lagrange polynomials, graphs, domain
 * constants) resident for many evaluations.  fixed_lagrange /
 * permutation_pk_lagrange (input_len rows each) and trace_omega are
 * validated and stored as the reference stores them. */
RustError midnight_sppark_bls12_381_numerator_context_create(
    void **out, size_t len, size_t input_len, const RawFr *fixed, const RawFr *fixed_lagrange,
    size_t fixed_cols, const RawFr *l0, const RawFr *l_last, const RawFr *l_active,
    const RawColumnRef *permutation_columns, size_t permutation_columns_len,
    const RawFr *permutation_pk_cosets, const RawFr *permutation_pk_lagrange,
    size_t permutation_pk_cosets_len, const uint32_t *trash_selector_fixed_indices,
    size_t trash_count, RawGraph custom_graph, const RawGraph *lookup_graphs,
    const RawGraph *lookup_input_graphs, const RawGraph *lookup_table_graphs,
    size_t lookup_count, const RawGraph *trash_graphs, RawFr omega, RawFr trace_omega,
    RawFr g_coset, RawFr delta_const, int32_t rot_scale, int32_t last_rotation,
    uint32_t chunk_len);

/* replaces bls12_381_g1_msm.cu:5662 */
void midnight_sppark_bls12_381_numerator_context_free(void *context);

/* replaces bls12_381_g1_msm.cu:5968: evaluates the numerator from resident
 * columns.  raw_columns is one device vector of total_columns * len values;
 * advice, instance, permutation products, lookup (product, input, table)
 * and trash cosets are column ranges of it.  delta_start = beta * g_coset.
 * *out receives a new device vector of len values. */
RustError midnight_sppark_bls12_381_numerator_context_eval(
    void **out, const void *raw_context, const void *raw_columns, size_t total_columns,
    size_t advice_offset, size_t advice_cols, size_t instance_offset, size_t instance_cols,
    const RawFr *challenges, size_t challenge_len, RawFr beta, RawFr gamma, RawFr theta,
    RawFr trash_challenge, RawFr y, size_t permutation_offset, size_t permutation_sets,
    size_t lookup_product_offset, size_t lookup_input_offset, size_t lookup_table_offset,
    size_t lookup_count, size_t trash_offset, size_t trash_count);

/* ---------------- grand products ---------------- */

/* replaces bls12_381_g1_msm.cu:4303: in place, values[0] = initial,
 * values[i] = initial * prod_{j < i} values_in[j] */
RustError midnight_sppark_bls12_381_fr_shifted_prefix_product(RawFr *values, RawFr initial, size_t len);

/* replaces bls12_381_g1_msm.cu:4342: permutation product columns of
 * column_count host columns (column-major, len rows each) with their
 * permuted identity values; sets of chunk_len columns chained through the
 * row len - blinding_factors - 1, the last blinding_factors rows of set s
 * taken from blind_values[s * blinding_factors ...]
 * (plonk/permutation/prover.rs:355-515).  *out receives a device vector of
 * ceil(column_count / chunk_len) * len values. */
RustError midnight_sppark_bls12_381_fr_permutation_products(
    void **out, const RawFr *columns, const RawFr *permutations, size_t len, size_t column_count,
    size_t chunk_len, RawFr beta, RawFr gamma, RawFr omega, const RawFr *delta_powers,
    const RawFr *blind_values, size_t blinding_factors);

/* replaces bls12_381_g1_msm.cu:4451: lookup product columns: z[0] = 1, z[i] = prod_{j < i}
 * (ci_j + beta)(ct_j + gamma) / ((pi_j + beta)(pt_j + gamma)) (zero denominator: zero fraction), the last
 * blinding_factors rows of lookup l from blind_values[l * blinding_factors ...]; inputs lookup-major
 * (lookup l at [l*len, (l+1)*len)) (plonk/lookup/prover.rs:1249-1340) */
RustError midnight_sppark_bls12_381_fr_lookup_products(
    void **out, const RawFr *compressed_inputs, const RawFr *compressed_tables, const RawFr *permuted_inputs,
    const RawFr *permuted_tables, size_t len, size_t lookup_count, RawFr beta, RawFr gamma,
    const RawFr *blind_values, size_t blinding_factors);
/* replaces bls12_381_g1_msm.cu:4545: the same from one device vector holding compressed inputs,
 * compressed tables, permuted inputs, permuted tables (4 * lookup_count * len values) */
RustError midnight_sppark_bls12_381_fr_lookup_products_device(void **out, const void *trace_lagrange, size_t len,
                                                              size_t lookup_count, RawFr beta, RawFr gamma,
                                                              const RawFr *blind_values, size_t blinding_factors);
/* replaces bls12_381_g1_msm.cu:6617: values sorted in place by canonical value (Fq's Ord) */
RustError midnight_sppark_bls12_381_fr_sort_canonical(RawFr *values, size_t len);
/* replaces bls12_381_g1_msm.cu:5139: *out = device vector [compressed inputs | compressed tables |
 * permuted inputs | permuted tables] (lookup-major blocks of len rows): the permuted columns of
 * permute_expression_pair (plonk/lookup/prover.rs:1573-1690) over the first usable_rows rows, then the
 * blinds (blind_count = len - usable_rows per lookup); code -1 when an input is not in its table */
RustError midnight_sppark_bls12_381_fr_lookup_permute_batch(void **out, const RawFr *compressed_inputs,
                                                            const RawFr *compressed_tables, size_t len,
                                                            size_t lookup_count, size_t usable_rows,
                                                            const RawFr *input_blinds, const RawFr *table_blinds,
                                                            size_t blind_count);
/* replaces bls12_381_g1_msm.cu:5849: the same with the compressed columns evaluated from a numerator
 * context's lookup input / table graphs (theta) on its Lagrange domain */
RustError midnight_sppark_bls12_381_trace_context_lookup_permute(
    void **out, const void *context, const void *advice, size_t advice_cols, const void *instance,
    size_t instance_cols, const RawFr *challenges, size_t challenge_len, RawFr theta, size_t usable_rows,
    const RawFr *input_blinds, const RawFr *table_blinds, size_t blind_count);
/* replaces bls12_381_g1_msm.cu:5667: each trash graph of a numerator context on the Lagrange domain
 * (input_len rows, rotations unscaled) from the context's fixed columns and resident advice / instance;
 * *out receives trash_count * input_len values */
RustError midnight_sppark_bls12_381_trace_context_trash(void **out, const void *context, const void *advice,
                                                        size_t advice_cols, const void *instance,
                                                        size_t instance_cols, const RawFr *challenges,
                                                        size_t challenge_len, RawFr trash_challenge);

/* replaces bls12_381_g1_msm.cu:5738: the same products from a numerator
 * context's Lagrange-form fixed and sigma columns (input_len rows, trace
 * omega) and resident advice / instance device vectors. */
RustError midnight_sppark_bls12_381_trace_context_permutation(
    void **out, const void *raw_context, const void *raw_advice, size_t advice_cols,
    const void *raw_instance, size_t instance_cols, RawFr beta, RawFr gamma,
    const RawFr *delta_powers, size_t delta_powers_len, const RawFr *blind_values,
    size_t blind_values_len, size_t blinding_factors);

/* ---------------- Poseidon hash chains (width 3, 8 full + 60 partial rounds, x^5) ---------------- */

/* replaces bls12_381_g1_msm.cu:4121 (lib.rs:1029): out[i] = previous, where each of the len hashes
 * permutes [previous, tail0, input_len], adds tail1 to lane 0 and permutes again, previous starting
 * at initial.  mds: 9 values (row-major), round_constants: 68 * 3; len == 0 is a no-op. */
RustError midnight_sppark_bls12_381_fr_poseidon_hash_chain(RawFr *out, size_t len, RawFr initial, RawFr tail0,
                                                          RawFr tail1, RawFr input_len, const RawFr *mds,
                                                          size_t mds_len, const RawFr *round_constants,
                                                          size_t round_constants_len);
/* replaces bls12_381_g1_msm.cu:4192 (lib.rs:1042): the same chain's circuit witness, 256 values per
 * hash (128 per permutation: per full round the S-box cubes and the next state, per batch of six
 * partial rounds five last-lane values and the state) */
RustError midnight_sppark_bls12_381_fr_poseidon_hash_chain_witness(RawFr *out, size_t len, RawFr initial,
                                                                  RawFr tail0, RawFr tail1, RawFr input_len,
                                                                  const RawFr *mds, size_t mds_len,
                                                                  const RawFr *round_constants,
                                                                  size_t round_constants_len);
/* replaces bls12_381_g1_msm.cu:2820 (lib.rs:854): a device vector of 8 * domain_len values (eight
 * advice columns, column-major) holding the chain witness placed by layout ((column << 8) | row, 128
 * entries) at rows 2 + 41 hash + 21 permutation, plus the input rows taken from metadata = {prefix[3],
 * binding, initial, tail0, tail1, input_len, round_constants[0..3]}; needs 2 + 41 hash_count <
 * domain_len, every layout column below 8 and every layout row inside the domain */
RustError midnight_sppark_bls12_381_fr_poseidon_advice_from_witness(void **out, const RawFr *witness,
                                                                   size_t hash_count, size_t domain_len,
                                                                   const RawFr *metadata, size_t metadata_len,
                                                                   const uint16_t *layout, size_t layout_len);

/* ---------------- backend extensions (no reference counterpart) ---------------- */

/* Number of elements of a device vector (for callers and tests that hold
 * only the opaque handle). */
size_t midnight_hip_fr_device_vec_len(const void *device_vec);

/* Batched ragged MSM over one resident point set: MSM j uses points
 * [0, sizes[j]) and scalars at scalar_offsets[j] of the device vector.  All
 * MSMs of the batch share each stage's launches.  For batch provers that
 * commit many polynomials of different sizes (many circuits of different k
 * in one launch). */
RustError midnight_hip_g1_msm_context_device_multi(RawG1Projective *out, void *context,
                                                   const void *device_vec,
                                                   const size_t *scalar_offsets,
                                                   const size_t *sizes, size_t count);

/* Batched ragged in-place inverse/forward NTTs over one device vector:
 * transform j covers 2^log_ns[j] values at offsets[j] (slices must not
 * overlap).  One launch per NTT pass for the whole set.  For batch provers
 * (many circuits of different k). */
RustError midnight_hip_fr_device_vec_ntt_multi(void *device_vec, const size_t *offsets,
                                               const uint32_t *log_ns, size_t count,
                                               bool inverse);

/* Batch plan: instances i < count, each an inverse NTT of the 2^log_ns[i]
 * values at offsets[i] of the device vector (in place) followed by the MSM
 * of those coefficients against points [0, 2^log_ns[i]) of the context.
 * A run does every instance: the ragged NTT (one launch per pass), then one
 * batched MSM.  flags: 1 = capture the run's device work once in a HIP graph
 * and replay it; 2 = autotune the MSM configuration (precomputed table with
 * chunk lengths 64/32, plain windows of the automatic width +-1) by timing
 * it on the batch at creation.  The reference picks MSM windows from a
 * per-k table (Bls12_381MsmConfig::for_k, midnight-sppark-msm/src/lib.rs:
 * 159-203) and runs instances one by one.  The plan refers to the context
 * and the device vector, which must outlive it; runs on the creating lane's
 * stream. */
RustError midnight_hip_batch_plan_create(void **out, void *msm_context, void *device_vec,
                                         const size_t *offsets, const uint32_t *log_ns, size_t count,
                                         uint32_t flags);
/* Runs the plan once; out[i] = instance i's MSM (Jacobian). */
RustError midnight_hip_batch_plan_run(void *plan, RawG1Projective *out);
/* The plan's MSM window, chunk length, whether it uses the precomputed
 * table, whether it replays a graph, and the tuned MSM time (ms, 0 if not
 * tuned). */
RustError midnight_hip_batch_plan_info(const void *plan, uint32_t *window_bits, uint32_t *chunk_len,
                                       uint32_t *precomputed, uint32_t *graph, float *tuned_ms);
void midnight_hip_batch_plan_free(void *plan);

/* In place: canonical Fr values -> Montgomery form (a -> a*R mod r).  For
 * callers that hold canonical scalars (synthetic workloads, wire formats). */
RustError midnight_hip_fr_device_vec_to_montgomery(void *device_vec);

/* MSM context over the points [t_i]G, where t_i are the (Montgomery) values
 * of a resident Fr vector, computed on the device by a fixed-base comb.
 * For SRS construction ([tau^i]G) and synthetic point sets, without a
 * host round trip of the points. */
RustError midnight_hip_g1_msm_context_create_from_multiples(void **out, const void *scalar_device_vec);
/* Copy the first n context points (RawG1Affine) to the host. */
RustError midnight_hip_g1_msm_context_points_download(RawG1Affine *out, const void *context, size_t n);
/* Statistics of the context's last MSM batch: bucket additions performed
 * (nonzero signed digits) and the window width used.  For throughput and
 * roofline reporting. */
/* The table window (or a caller config's plain window), accumulation chunk
 * bound and autotuned MSM time in ms (0 when not tuned) of a context.
 * Contexts made without a caller config and of 2^16 points or more are
 * autotuned when created (the table window of the size rule and one bit
 * either side, chunks of 64 and 32 entries, timed on a batch of four MSMs
 * of uniform scalars); with a config, a nonzero window_bits is used as given (table
 * windows capped at 16).  MSM results never depend on the choice. */
RustError midnight_hip_g1_msm_context_tuning(const void *context, uint32_t *window_bits, uint32_t *chunk_len,
                                            float *tuned_ms);
RustError midnight_hip_g1_msm_context_last_stats(const void *context, uint64_t *entries,
                                                 uint32_t *window_bits);

/* ---- proof pipeline (create_proof of plonk/prover.rs:1019-1076 for circuits
 * made of custom gates, a copy permutation, lookups, trashcans and
 * non-committed instance columns).  For provers that keep a circuit's key
 * and the SRS resident and want whole proofs from one call. ---- */

/* Proving key.  srs_context: an MSM context whose first srs_points points
 * are [tau^i]G (>= 2^k); not owned, must outlive the key.  fixed: fixed_cols
 * columns of 2^k Lagrange values.  advice_queries / fixed_queries: pairs
 * (column, rotation as int32 bits), in the constraint system's query order.
 * perm_columns: the permuted columns in order; perm_mapping: for permuted
 * column j and row i, the pair (j', i') of the next cell of its cycle
 * (itself when not copied).  gate: the custom-gate graph (Horner in y).
 * description: bytes hashed into the verifying-key representative. */
RustError midnight_hip_plonk_keygen(void **out, uint32_t k, void *srs_context, size_t srs_points,
                                    uint32_t degree, const RawFr *fixed, size_t fixed_cols,
                                    size_t advice_cols, const uint32_t *advice_queries,
                                    size_t advice_query_count, const uint32_t *fixed_queries,
                                    size_t fixed_query_count, const RawColumnRef *perm_columns,
                                    size_t perm_cols, const uint32_t *perm_mapping, RawGraph gate,
                                    const char *description);
/* The arguments of a key beyond gates and the copy permutation.
 * instance_cols: instance columns (compute_instances,
 * plonk/prover.rs:1237-1330), the first committed_instances of them
 * committed (nb_committed_instances: the commitment of each such column's
 * Lagrange polynomial goes into the transcript, its instance queries are
 * evaluated after x and opened; instance_queries lists (column, rotation
 * as int32 bits) pairs in the constraint system's order), the others
 * non-committed (their lengths and values go into the transcript; neither
 * committed nor opened).  Per lookup
 * (plonk/lookup/prover.rs): the compression of its input expressions and of
 * its table expressions (Horner in theta from zero, evaluation.rs:2743-2766)
 * and the numerator's lookup graph (input + beta)(table + gamma)
 * (evaluation.rs:2768-2797).  Per trashcan (plonk/trash/prover.rs): the graph
 * of its constraint expressions folded in the trash challenge and the fixed
 * column of its selector.  Graphs read advice at rotations the key queries. */
typedef struct RawPlonkArguments {
    size_t instance_cols;
    size_t committed_instances;
    const uint32_t *instance_queries;
    size_t instance_query_count;
    size_t lookup_count;
    const RawGraph *lookup_input_graphs;
    const RawGraph *lookup_table_graphs;
    const RawGraph *lookup_graphs;
    size_t trash_count;
    const RawGraph *trash_graphs;
    const uint32_t *trash_selectors;
} RawPlonkArguments;
/* midnight_hip_plonk_keygen for a circuit with the arguments above (null:
 * none).  Permuted columns may name instance columns (kind 2). */
RustError midnight_hip_plonk_keygen_with_arguments(
    void **out, uint32_t k, void *srs_context, size_t srs_points, uint32_t degree, const RawFr *fixed,
    size_t fixed_cols, size_t advice_cols, const uint32_t *advice_queries, size_t advice_query_count,
    const uint32_t *fixed_queries, size_t fixed_query_count, const RawColumnRef *perm_columns, size_t perm_cols,
    const uint32_t *perm_mapping, RawGraph gate, const char *description, const RawPlonkArguments *arguments);
void midnight_hip_plonk_key_free(void *pk);
/* Verifying data: fixed and permutation commitments, transcript representative. */
RustError midnight_hip_plonk_vk(const void *pk, RawG1Projective *fixed_commitments,
                                RawG1Projective *permutation_commitments, RawFr *transcript_repr);
/* One proof: advice = advice_cols columns of 2^k Lagrange values (rows from
 * 2^k - blinding - 1 on are overwritten with blinding values), blinding
 * drawn from ChaCha20Rng::seed_from_u64(rng_seed).  *proof_len receives the
 * proof size; a null proof or a proof_capacity below it is rejected before
 * any work is done (the size is fixed by the key). */
RustError midnight_hip_plonk_prove(void *pk, const RawFr *advice, size_t advice_len, uint64_t rng_seed,
                                   uint8_t *proof, size_t proof_capacity, size_t *proof_len);
/* The same with the proof's instance values: instance_cols (the key's count)
 * columns of instance_lens[c] values each, concatenated in column order; a
 * column may hold at most 2^k - blinding - 1 values (Error::InstanceTooLarge
 * of compute_instances), the rest of its rows are zero.  A lookup input
 * missing from its table fails the proof (code -1) as
 * permute_expression_pair does. */
RustError midnight_hip_plonk_prove_with_instances(void *pk, const RawFr *advice, size_t advice_len,
                                                  const RawFr *instance, const size_t *instance_lens,
                                                  size_t instance_cols, uint64_t rng_seed, uint8_t *proof,
                                                  size_t proof_capacity, size_t *proof_len);
/* Commitment stages from HIP graphs (default on): on each lane, the first
 * proof of the key captures its commitment stages (inverse NTT + MSM batch of
 * the advice, the permutation products; the MSMs of the random polynomial,
 * the quotient pieces and the opening polynomials) and every later proof on
 * that lane replays them; the lane keeps its polynomial arena and MSM
 * engines for the key's lifetime.  Off: every stage is launched eagerly.
 * Proof bytes are the same either way.  Not to be called while a proof of
 * the key runs. */
RustError midnight_hip_plonk_set_graphs(void *pk, bool enable);
/* The length in bytes of every proof of this key. */
RustError midnight_hip_plonk_proof_size(const void *pk, size_t *proof_len);
/* Thread count of the reference prover whose draws proofs of this key
 * reproduce (default 1): the vanishing argument's random polynomial takes one
 * ChaCha20Rng seed per chunk of that many workers (vanishing/prover.rs:41-75,
 * split as utils/arithmetic.rs:138-190 parallelize), which shifts every later
 * draw of the prover's RNG. */
RustError midnight_hip_plonk_set_rng_threads(void *pk, uint32_t threads);
/* vanishing::Argument::sample_random (vanishing/prover.rs:41-75) once its
 * seeds are drawn: a new device vector of 2^log_n Fr, chunk c of the
 * parallelize split of 2^log_n rows over `threads` workers holding the first
 * Fq::random values of ChaCha20Rng::from_seed(seeds + 32 c); seed_count must
 * equal the chunk count, min(threads, 2^log_n). */
RustError midnight_hip_fr_sample_random_poly(void **out, uint32_t log_n, size_t threads, const uint8_t *seeds,
                                             size_t seed_count);

/* KZG parameter files (ParamsKZG::read_custom, poly/kzg/params.rs:187-262):
 * u32 k || 2^k G1 g || 2^k G1 g_lagrange || G2 || [tau]G2, parsed and checked
 * on the device into two resident MSM contexts.  format follows SerdeFormat
 * (utils/helpers.rs:16-28): 0 Processed (compressed G1, subgroup-checked),
 * 1 RawBytes (uncompressed G1, checked below p, on the curve and in the
 * subgroup), 2 RawBytesUnchecked (uncompressed, coordinates below p only).
 * An invalid point fails the load.  The G2 elements are kept as read. */
RustError midnight_hip_srs_load(void **out, const uint8_t *bytes, size_t len, uint32_t format);
RustError midnight_hip_srs_load_file(void **out, const char *path, uint32_t format);
/* k, and the MSM contexts of g and g_lagrange (owned by the SRS: use them
 * with every midnight_sppark_bls12_381_g1_msm_context_* entry point, never
 * free them; committing Lagrange values against g_lagrange is
 * KZGCommitmentScheme::commit_lagrange, poly/kzg/mod.rs:1464-1536). */
RustError midnight_hip_srs_info(const void *srs, uint32_t *k, void **g_context, void **g_lagrange_context);
/* The file's G2 || [tau]G2 bytes. */
RustError midnight_hip_srs_g2(const void *srs, uint8_t *out, size_t capacity, size_t *len);
void midnight_hip_srs_free(void *srs);

/* Proof server: `slots` worker threads (lanes 1..slots, at most 8) prove
 * submitted requests concurrently on the GPU; a request is admitted when a
 * slot is free and its HBM footprint fits in hbm_budget_bytes beside the
 * running proofs.  Keys passed to submit must outlive their requests.  For
 * serving processes (proof-server's worker pool) that batch concurrent
 * requests of different circuit sizes. */
RustError midnight_hip_proof_server_create(void **out, uint32_t slots, uint64_t hbm_budget_bytes);
/* Queues one proof (the advice is copied); *job_id identifies it. */
RustError midnight_hip_proof_server_submit(void *server, void *pk, const RawFr *advice, size_t advice_len,
                                           uint64_t seed, uint64_t *job_id);
/* The same with the proof's instance values (instance_cols columns of
 * instance_lens[c] values, concatenated, as midnight_hip_plonk_prove_with_instances
 * takes them); the values are copied at submission. */
RustError midnight_hip_proof_server_submit_with_instances(void *server, void *pk, const RawFr *advice,
                                                          size_t advice_len, const RawFr *instance,
                                                          const size_t *instance_lens, size_t instance_cols,
                                                          uint64_t seed, uint64_t *job_id);
/* Waits for a job: proof bytes, and its time queued and proving (ms).  A
 * null or too small proof buffer is an error that sets *proof_len and keeps
 * the finished job, so the caller can wait again with a large enough one. */
RustError midnight_hip_proof_server_wait(void *server, uint64_t job_id, uint8_t *proof, size_t capacity,
                                         size_t *proof_len, double *queue_ms, double *prove_ms);
/* Finishes queued and running proofs, then stops the server. */
void midnight_hip_proof_server_free(void *server);

/* Measured throughput of the integer multiply-add the field code is built
 * on (v_mad_u64_u32), in operations per second over the whole device, and
 * the device clock/CU count used; for roofline reporting. */
RustError midnight_hip_valu_peak(double *mad_u64_per_s, double *clock_hz, int *num_cus);
/* One device-to-device copy of `bytes` (multiple of 16) by a plain copy
 * kernel: read + write bytes per second.  Also the launch of known size that
 * hardware byte counters are scaled against. */
RustError midnight_hip_copy_bandwidth(size_t bytes, double *bytes_per_s);
/* Whole Fp / Fr Montgomery products per second (blocks = 0: whole device;
 * blocks = 1: one wave, i.e. the latency of a dependent chain). */
RustError midnight_hip_field_mul_rate(int blocks, double *fp_per_s, double *fr_per_s);
/* Latency in seconds of one dependent radix-2^29 Fr product on a single
 * wave (the product the Poseidon chain kernel issues). */
RustError midnight_hip_fr29_product_latency(double *seconds);
/* Per-kernel timing with HIP events on the launching stream (off by
 * default).  enable(1) clears the registry; read() returns 1 and the total
 * milliseconds and count for a region name ("msm_accumulate", "msm_total",
 * ...), 0 if the region has not run. */
void midnight_hip_profile_enable(int on);
int midnight_hip_profile_read(const char *name, double *total_ms, uint64_t *count);

#ifdef __cplusplus
}
#endif
#endif
