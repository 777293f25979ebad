// Resident MSM point set behind the reference's opaque context handle
// (msm_context_t, bls12_381_g1_msm.cu:154-329): the points stay in HBM for
// the context's lifetime and every MSM of the context reuses the engine's
// scratch buffers.
#pragma once
#include <memory>
#include <mutex>

#include "msm.hpp"
#include "runtime.hpp"

namespace mh {

struct MsmContext {
    uint32_t lane = 0;
    RawMsmConfig config{};
    size_t npoints = 0;
    DevBuf<uint32_t> points;   // npoints rows of RawG1Affine (24 u32)
    // per-window copies [2^(c w)] P_i (w < pc_nwin), built once per context
    DevBuf<uint32_t> table;
    uint32_t pc_c = 0, pc_nwin = 0;
    // A narrow-window table [2^(sm_c w)] P_i over the first sm_n points, for
    // MSMs too short for the main table's 2^(pc_c - 1) buckets: its one
    // bucket set stays small and the host finishes with sm_c doublings, where
    // plain windows leave a doubling chain over all 256 bits.
    DevBuf<uint32_t> small_table;
    uint32_t sm_c = 0, sm_nwin = 0;
    size_t sm_n = 0;
    uint32_t window_bits = 0;  // a caller config's window (0: autotuned table / automatic plain windows)
    uint32_t chunk_len = 64;   // accumulation chunk bound of this context's MSMs
    float tuned_ms = 0;        // autotuned MSM time at creation (0: not tuned)
    // One engine (scratch buffers, statistics) per lane, so that MSMs of
    // different lanes over the same points run concurrently; calls of one
    // lane take turns on its engine.
    struct LaneEngine {
        std::mutex mu;
        MsmEngine engine;
    };
    std::mutex mu;   // guards creation of lane engines
    std::unique_ptr<LaneEngine> engines[kLaneCount];
    LaneEngine &lane_engine(uint32_t lane)
    {
        std::lock_guard<std::mutex> g(mu);
        if (!engines[lane]) engines[lane] = std::make_unique<LaneEngine>();
        return *engines[lane];
    }
};

RawMsmConfig default_msm_config(size_t npoints);
std::unique_ptr<MsmContext> make_context(size_t npoints);
// Builds the per-window table of a context whose points are on the device
// (no-op for small point sets, which use plain windows): the config's
// window_bits when set, else autotuned for contexts of 2^16 points or more.
void build_precomp(MsmContext &ctx, hipStream_t st);
// The precomputed table for a batch whose longest MSM has max_n points and
// whose points all lie below max_end: the main table when max_n reaches
// half its bucket count, else the small table when it covers max_end.
// False: plain windows.
bool context_table(const MsmContext &ctx, size_t max_n, size_t max_end, PrecompTable &pc);
void generator_multiples(uint32_t *out_points, const Fr *scalars, size_t n, hipStream_t st);

}  // namespace mh
