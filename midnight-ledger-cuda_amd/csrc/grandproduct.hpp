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

// Lookup product columns, lookup-major [count][len]: z[0] = 1,
// z[k] = prod_{j < k} (ci_j + beta)(ct_j + gamma) / ((pi_j + beta)(pt_j + gamma)),
// the last `blinding` rows of lookup l from blinds[l * blinding ...].
void lookup_products_device(Fr *out, const Fr *ci, const Fr *ct, const Fr *pi, const Fr *pt, size_t len,
                            size_t count, const HostFr &beta, const HostFr &gamma, const Fr *blinds, size_t blinding,
                            hipStream_t st);

// out[0] = initial, out[i] = initial * prod_{j < i} in[j] (in and out may alias).
void shifted_prefix_product_device(Fr *out, const Fr *in, size_t len, const HostFr &initial, hipStream_t st);

}  // namespace mh
