// Coefficient-form polynomial operations on the device (evaluation at
// points, linear combinations, division by X - a).
#pragma once
#include <vector>

#include "fields.hpp"
#include "host_fr.hpp"
#include "runtime.hpp"

namespace mh {

// out[poly * npts + pt] = polys[poly](points[pt]); count polys of len
// coefficients each, contiguous.
void poly_eval_many(HostFr *out, const Fr *polys, size_t len, size_t count, const HostFr *points, size_t npts,
                    hipStream_t st);
// out = constant + sum_j scalars[j] * polys[j] (device pointers, len each)
void poly_lincomb(Fr *out, const std::vector<const Fr *> &polys, const std::vector<HostFr> &scalars, size_t len,
                  const HostFr &constant, hipStream_t st);
// q = (p - p(a)) / (X - a): q[0 .. len-2], q[len-1] = 0; q may alias p
void poly_kate_divide(Fr *q, const Fr *p, size_t len, const HostFr &a, hipStream_t st);
// out[q] = polys[qpoly[q]](points[qpt[q]]), polys of len coefficients each
void poly_eval_queries(HostFr *out, const std::vector<const Fr *> &polys, size_t len, const uint32_t *qpoly,
                       const uint32_t *qpt, size_t nq, const HostFr *points, size_t npts, hipStream_t st);
// out[q * len + i] = sum_{t in [offsets[q], offsets[q+1])} coefficients[t] * polys[poly_idx[t]][i]
void poly_q_combine(Fr *out, const std::vector<const Fr *> &polys, size_t len, const uint32_t *offsets,
                    size_t q_count, const uint32_t *poly_idx, const HostFr *coefficients, size_t items,
                    hipStream_t st);

}  // namespace mh
