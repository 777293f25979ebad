// Host interface of the batched Pippenger engine (msm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <memory>
#include <vector>

#include "fields.hpp"
#include "runtime.hpp"

namespace mh {

class ProfileScope;

// One MSM of a batch: n scalars (device, Montgomery) against points
// [pt_off, pt_off + n) of the engine's point table.
struct MsmJob {
    const Fr *scalars;
    size_t pt_off;
    size_t n;
};

// Signed c-bit windows of a scalar: every scalar above (r - 1) / 2 is
// replaced by r - s with its point negated, so scalars are below 2^254 and
// ceil(255 / c) windows hold them with the last carry (one window fewer than
// ceil(256 / c) at c = 15 and 17).
__host__ __device__ constexpr uint32_t msm_windows(uint32_t c) { return (255 + c - 1) / c; }

// Whether the top window of width c holds at least four bits of a scalar
// (scalars are below 2^254).  With a precomputed table every window feeds
// one bucket set, so a top window of one or two bits (c = 9, 11, 12, 14)
// sends all of its digits to the lowest few buckets, whose long chunk
// trees then set the latency of the whole MSM.
__host__ __device__ constexpr bool full_top_window(uint32_t c) { return 254 >= c * (msm_windows(c) - 1) + 4; }

// Per-window copies of a point set: row w * n + i holds [2^(c w)] P_i
// (RawG1Affine), w < nwin = msm_windows(c).  With it every window of an MSM
// feeds one shared bucket set, and the window fold disappears.
struct PrecompTable {
    const uint32_t *table;
    size_t n;
    uint32_t c, nwin;
};

// window width the engine picks for plain windows over MSMs of n points
uint32_t msm_auto_window(uint64_t n);

class MsmEngine {
public:
    MsmEngine();
    ~MsmEngine();
    // d_points: RawG1Affine rows (24 u32 each).  window_bits 0 = automatic.
    // With `pc`, the point rows come from the precomputed table instead.
    void run(const uint32_t *d_points, size_t npoints_total, const std::vector<MsmJob> &jobs, uint32_t window_bits,
             RawG1Projective *out, hipStream_t st, const PrecompTable *pc = nullptr);

    // run() in three parts, for callers that capture the device work in a
    // graph: prepare() sizes the buffers for the worst case of the batch and
    // uploads its descriptors (stream-ordered copies, no kernels); enqueue()
    // only launches kernels and copies the (window, plane) sums to pinned
    // host memory -- no allocation, no synchronisation, no memset -- and may
    // be replayed; finish() runs the doubling chains once the stream has
    // synchronised.  The buffers stay valid until the next prepare().
    void prepare(const uint32_t *d_points, size_t npoints_total, const std::vector<MsmJob> &jobs,
                 uint32_t window_bits, hipStream_t st, const PrecompTable *pc = nullptr);
    void enqueue(hipStream_t st, ProfileScope *prof);
    void finish(RawG1Projective *out);
    // enqueue() in two parts: through the accumulation, then the bucket
    // tree with `levels` levels (at most the batch's worst case; fewer is
    // still exact, the last pass sums what is left), the planes and the
    // copy-out.  The chunk count of the fullest bucket is totals[3] on the
    // device after enqueue_front().
    void enqueue_front(hipStream_t st, ProfileScope *prof);
    void enqueue_back(hipStream_t st, uint32_t levels, ProfileScope *prof);

    // Upper bounds of the serial work per lane.  A point addition takes the
    // same time per lane at any occupancy, so batches too small to fill the
    // GPU get shorter chunks and groups (down to 4): more lanes, a shorter
    // serial chain.
    uint32_t chunk_len = 64;      // entries per accumulation lane
    uint32_t plane_group = 16;    // buckets per lane in the first plane pass
    uint32_t plane_group2 = 8;    // partials per lane in the later plane passes
    // Reduction passes of fewer sums than this give each sum a quad of
    // lanes (four products per step of the addition instead of one): about
    // a third of the latency per addition for a third more issue, a gain
    // while the pass leaves SIMDs without a wave (0 = never).
    uint64_t quad_below = 3 * 16384;
    // Reduction passes of fewer sums than this (a batch of five MSMs over
    // a 4096-bucket table included) give each sum four quads
    // (each adds every fourth partial, then the quads' sums are added
    // pairwise): about a quarter of a quad's chain for a third more issue.
    uint64_t group16_below = 24576;

    // statistics of the last run (for throughput / roofline reporting)
    uint64_t last_entries = 0;    // nonzero signed digits = bucket additions
    uint32_t last_window = 0;

private:
    struct Buffers;
    struct State;
    std::unique_ptr<Buffers> buf_;
    std::unique_ptr<State> st_;
};

}  // namespace mh
