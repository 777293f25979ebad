// Device ChaCha20Rng draws of field elements.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

#include "fields.hpp"

namespace mh {

// out[0..n) = the first n Fq::random values of ChaCha20Rng::from_seed(seed),
// Montgomery form, written on stream st.  Synchronizes st once (after a
// short counting pass) to check that enough candidates were accepted.
void random_fr_device(Fr *out, size_t n, const uint8_t seed[32], hipStream_t st);

// Several streams at once: out[offsets[s] .. offsets[s] + lens[s]) = the
// first lens[s] Fq::random values of ChaCha20Rng::from_seed(seeds + 32 s).
// Synchronizes st once, like random_fr_device.
void random_fr_streams_device(Fr *out, const size_t *offsets, const size_t *lens, const uint8_t *seeds,
                              size_t count, hipStream_t st);

// The chunks parallelize (utils/arithmetic.rs:138-190) cuts n rows into for
// `threads` workers: n % threads chunks of n / threads + 1 rows first, then
// chunks of n / threads rows (none when that is 0), in row order.
void parallelize_chunks(size_t n, size_t threads, std::vector<size_t> &offsets, std::vector<size_t> &lens);

}  // namespace mh
