// Host interface of the element-wise Fr kernels (frops.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "fields.hpp"

namespace mh {

void fr_mul_scalar(Fr *v, const Fr &s, size_t n, hipStream_t st);
void fr_mul_zeta(Fr *v, const Fr &first, const Fr &second, size_t n, hipStream_t st);
void fr_mul_repeating(Fr *v, const Fr *factors, size_t m, size_t n, hipStream_t st);
void fr_add_mask(Fr *out, const Fr *in, const Fr *mask, size_t m, size_t n, hipStream_t st);
void fr_quotient_pieces(Fr *out, const Fr *quotient, size_t poly_len, size_t pieces, const Fr *blinds,
                        hipStream_t st);
// the same with the pieces - 1 blinds from host memory (at most
// kMaxArgBlinds), passed by value: one launch, no upload
constexpr size_t kMaxArgBlinds = 8;
void fr_quotient_pieces(Fr *out, const Fr *quotient, size_t poly_len, size_t pieces, const Fr *host_blinds,
                        size_t nblinds, hipStream_t st);
void fr_fill(Fr *v, const Fr &x, size_t n, hipStream_t st);

// in place: v[i] = 1 / v[i], zeros stay zero
void fr_batch_invert(Fr *v, size_t n, hipStream_t st);

}  // namespace mh
