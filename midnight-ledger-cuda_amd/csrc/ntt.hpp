// Host interface of the Fr NTT passes (ntt.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>

#include "fields.hpp"
#include "host_fr.hpp"

namespace mh {

struct NttJob {
    const Fr *in = nullptr;      // count polys, in_pstride apart
    Fr *out = nullptr;           // count polys of 2^log_n, out_pstride apart (may equal in)
    Fr *scratch = nullptr;       // ntt_scratch_elems(log_n, count) elements
    uint64_t in_pstride = 0, out_pstride = 0;
    uint32_t log_n = 0;
    uint32_t count = 1;
    bool inverse = false;        // inverse includes the 1/n factor
    uint32_t pre = 0;            // 0 none, 1 zeta + zero pad past in_len, 2 repeating factors
    uint64_t in_len = ~0ull;
    const Fr *factors = nullptr;
    uint64_t factor_len = 1;
    bool post_zeta = false;      // multiply by scale * {1, first, second}[i % 3]
    HostFr scale = HostFr::one();
    HostFr first = HostFr::one(), second = HostFr::one();
    bool post_mul = false;       // also multiply every output by `mul`
    HostFr mul = HostFr::one();
    // Two-pass transforms read their four-step twiddles from a table per
    // size and direction, built (allocated) on first use; false forms them
    // from two small static tables instead (no allocation).
    bool mid_table = true;
};

void ntt_run(const NttJob &job, hipStream_t st);
// Builds the four-step twiddle tables of 2^log_n (both directions) now, so
// that transforms captured in a graph later find them.
void ntt_warm_tables(uint32_t log_n, hipStream_t st);
size_t ntt_scratch_elems(uint32_t log_n, size_t count);

// In-place transforms of mixed sizes (polynomial i: 2^log_ns[i] values at
// base + offsets[i]) with one launch per pass for the whole set: every
// transform's first pass in one launch, the second passes of those above
// 2^11 in another.  prepare() uploads the pass descriptors and sizes the
// set's own scratch (stream-ordered); enqueue() only launches, so it can be
// captured in a graph and replayed.  The buffers stay valid until the next
// prepare().
class RaggedNtt {
public:
    RaggedNtt();
    ~RaggedNtt();
    void prepare(Fr *base, const size_t *offsets, const uint32_t *log_ns, size_t count, bool inverse,
                 hipStream_t st);
    void enqueue(hipStream_t st);

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace mh
