// Device ChaCha20Rng draws of field elements.
#pragma once
#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

#include "fields.hpp"

namespace mh {

// out[0..n) = the first n Fq::random values of ChaCha20Rng::from_seed(seed),
// Montgomery form, written on stream st.  Synchronizes st once (after a
// short counting pass) to check that enough candidates were accepted.
void random_fr_device(Fr *out, size_t n, const uint8_t seed[32], hipStream_t st);

}  // namespace mh
