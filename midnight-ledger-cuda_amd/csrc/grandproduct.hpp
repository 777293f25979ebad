// Grand products on the device: multiplicative scans and the permutation
// argument's product columns.
#pragma once
#include "fields.hpp"
#include "host_fr.hpp"
#include "runtime.hpp"

namespace mh {

// Where the permutation's columns live on the device (column c of a kind
// at base + c * len); refs is a device array of column_count entries.
struct PermColumnsView {
    const Fr *fixed = nullptr, *advice = nullptr, *instance = nullptr;
    uint32_t fixed_cols = 0, advice_cols = 0, instance_cols = 0;
    const RawColumnRef *refs = nullptr;
};

// Grand products of column_count columns in sets of chunk_len
// (permutation/prover.rs:355-515 semantics): out receives
// ceil(column_count / chunk_len) * len values; sigmas holds the permuted
// identity values (column_count * len); delta_powers[j] is the identity
// multiplier of column j; blinds (device) the sets' blinding rows.
void permutation_products_device(Fr *out, size_t len, size_t column_count, size_t chunk_len,
                                 const PermColumnsView &cols, const Fr *sigmas, const HostFr *delta_powers,
                                 const HostFr &beta, const HostFr &gamma, const HostFr &omega, const Fr *blinds,
                                 size_t blinding, hipStream_t st);

// out[0] = initial, out[i] = initial * prod_{j < i} in[j] (in and out may alias).
void shifted_prefix_product_device(Fr *out, const Fr *in, size_t len, const HostFr &initial, hipStream_t st);

}  // namespace mh
