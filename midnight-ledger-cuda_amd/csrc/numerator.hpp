// Host interface of the quotient numerator (numerator.hip) beyond the C ABI.
#pragma once
#include <cstddef>

#include "runtime.hpp"

namespace mh {

// midnight_sppark_bls12_381_numerator_context_eval; with columns_scaled the
// column vector holds its values in the scaled form v 2^261 mod r (8 x 32
// bits) instead of the usual v 2^256 mod r, and is read without conversion;
// with out_scaled the result is written in that form too (32 times the
// usual form's value).
RustError numerator_context_eval(void **out, const void *raw_context, const void *raw_columns, size_t total_columns,
                                 size_t advice_offset, size_t advice_cols, size_t instance_offset,
                                 size_t instance_cols, const RawFr *challenges, size_t challenge_len, RawFr beta,
                                 RawFr gamma, RawFr theta, RawFr trash_challenge, RawFr y, size_t permutation_offset,
                                 size_t permutation_sets, size_t lookup_product_offset, size_t lookup_input_offset,
                                 size_t lookup_table_offset, size_t lookup_count, size_t trash_offset,
                                 size_t trash_count, bool columns_scaled, bool out_scaled = false);

}  // namespace mh
