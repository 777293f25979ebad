// Lookup-argument permutation and canonical sorting on the device.
#pragma once
#include <cstddef>

#include <hip/hip_runtime.h>

#include "fields.hpp"
#include "runtime.hpp"

namespace mh {

// out = v sorted by canonical value (ascending); out may not alias v.
void fr_sort_canonical_device(Fr *out, const Fr *v, size_t n, hipStream_t st);

// trace: 4 * lookup_count * len values, lookup-major blocks [compressed
// inputs | compressed tables | permuted inputs | permuted tables]; fills the
// permuted blocks from the compressed ones over the first `usable` rows and
// the blinds (device, lookup_count * blind_count each) after them.  Returns
// false when an input value is not in its table.
bool lookup_permute_device(Fr *trace, size_t len, size_t lookup_count, size_t usable, const Fr *input_blinds,
                           const Fr *table_blinds, size_t blind_count, hipStream_t st);

}  // namespace mh
