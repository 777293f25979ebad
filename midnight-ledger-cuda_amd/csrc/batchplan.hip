// Batched iNTT + MSM plans: many independent instances of uneven sizes
// (instance i: an inverse NTT of 2^k_i values in place, then the MSM of the
// resulting coefficients against the first 2^k_i points of a resident
// context), executed as one launch set for the whole batch -- the ragged NTT
// (one launch per pass), then one batched Pippenger over every instance --
// and captured once in a HIP graph that each run replays.
//
// The reference runs such work one polynomial at a time (fr_ntt,
// bls12_381_g1_msm.cu:6488-6502, then context_device_batch per MSM) and
// picks its MSM window from a hand-measured per-k table
// (Bls12_381MsmConfig::for_k, midnight-sppark-msm/src/lib.rs:159-203).
// Here the plan times the candidate MSM configurations on the batch itself
// when it is created (the precomputed per-window table with chunk lengths
// 64 and 32, and plain windows of the automatic width and one bit either
// side) and keeps the fastest.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "devvec.hpp"
#include "msm.hpp"
#include "msm_context.hpp"
#include "ntt.hpp"
#include "profile.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

enum : uint32_t { kPlanGraph = 1, kPlanAutotune = 2 };

struct Candidate {
    bool precomputed;
    uint32_t window;      // plain windows: width (0 = automatic)
    uint32_t chunk_len;
    bool small = false;   // precomputed: the context's small table
};

}  // namespace

struct BatchPlan {
    uint32_t lane = 0;
    MsmContext *ctx = nullptr;
    DeviceVec *dv = nullptr;
    std::vector<size_t> offsets;
    std::vector<uint32_t> log_ns;
    std::vector<MsmJob> jobs;
    RaggedNtt ntt;
    MsmEngine engine;
    Candidate chosen{};
    float tuned_ms = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::mutex mu;

    ~BatchPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }

    void prepare_msm(const Candidate &c, hipStream_t st)
    {
        engine.chunk_len = c.chunk_len;
        if (c.precomputed) {
            const PrecompTable pc = c.small ? PrecompTable{ctx->small_table.get(), ctx->sm_n, ctx->sm_c, ctx->sm_nwin}
                                            : PrecompTable{ctx->table.get(), ctx->npoints, ctx->pc_c, ctx->pc_nwin};
            engine.prepare(ctx->points.get(), ctx->npoints, jobs, 0, st, &pc);
        } else {
            engine.prepare(ctx->points.get(), ctx->npoints, jobs, c.window, st, nullptr);
        }
    }

    // milliseconds of the MSM part with candidate c (best of two runs)
    float time_msm(const Candidate &c, hipStream_t st)
    {
        prepare_msm(c, st);
        hipEvent_t a, b;
        MH_HIP(hipEventCreate(&a));
        MH_HIP(hipEventCreate(&b));
        float best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            MH_HIP(hipEventRecord(a, st));
            engine.enqueue(st, nullptr);
            MH_HIP(hipEventRecord(b, st));
            MH_HIP(hipEventSynchronize(b));
            float ms = 0;
            MH_HIP(hipEventElapsedTime(&ms, a, b));
            best = std::min(best, ms);
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return best;
    }
};

}  // namespace mh

using namespace mh;

extern "C" RustError midnight_hip_batch_plan_create(void **out, void *msm_context, void *device_vec,
                                                    const size_t *offsets, const uint32_t *log_ns, size_t count,
                                                    uint32_t flags)
{
    if (!out) return make_error((int)hipErrorInvalidValue, "out is null");
    *out = nullptr;
    if (!msm_context || !device_vec || count == 0 || !offsets || !log_ns)
        return make_error((int)hipErrorInvalidValue, "invalid batch plan arguments");
    auto *ctx = static_cast<MsmContext *>(msm_context);
    auto *dv = static_cast<DeviceVec *>(device_vec);
    std::vector<std::pair<size_t, size_t>> spans(count);
    for (size_t i = 0; i < count; ++i) {
        if (log_ns[i] > 22) return make_error((int)hipErrorInvalidValue, "batch instance above 2^22");
        const size_t n = size_t(1) << log_ns[i];
        if (offsets[i] > dv->len || n > dv->len - offsets[i])
            return make_error((int)hipErrorInvalidValue, "batch instance exceeds the device vector");
        if (n > ctx->npoints) return make_error((int)hipErrorInvalidValue, "batch instance exceeds the MSM context");
        spans[i] = {offsets[i], n};
    }
    std::sort(spans.begin(), spans.end());
    for (size_t i = 1; i < count; ++i)
        if (spans[i].first < spans[i - 1].first + spans[i - 1].second)
            return make_error((int)hipErrorInvalidValue, "batch instances overlap");
    MH_TRY
    auto plan = std::make_unique<BatchPlan>();
    plan->lane = current_lane();
    plan->ctx = ctx;
    plan->dv = dv;
    plan->offsets.assign(offsets, offsets + count);
    plan->log_ns.assign(log_ns, log_ns + count);
    hipStream_t st = lane_stream(plan->lane);
    size_t max_n = 0;
    plan->jobs.resize(count);
    for (size_t i = 0; i < count; ++i) {
        const size_t n = size_t(1) << log_ns[i];
        plan->jobs[i] = MsmJob{dv->values.get() + offsets[i], 0, n};
        max_n = std::max(max_n, n);
    }
    plan->ntt.prepare(dv->values.get(), plan->offsets.data(), plan->log_ns.data(), count, true, st);

    // the MSM configuration: timed candidates, or the engine's own choice
    std::vector<Candidate> cands;
    const bool has_pc = ctx->pc_nwin != 0;
    size_t max_end = 0;
    for (const MsmJob &j : plan->jobs) max_end = std::max<size_t>(max_end, j.pt_off + j.n);
    const bool has_small = ctx->sm_nwin != 0 && max_end <= ctx->sm_n;
    PrecompTable pc;
    if (flags & kPlanAutotune) {
        if (has_pc) {
            cands.push_back({true, 0, 64});
            cands.push_back({true, 0, 32});
        }
        if (has_small) cands.push_back({true, 0, 64, true});
        const uint32_t w = msm_auto_window(max_n);
        for (uint32_t c : {w - 1, w, w + 1})
            if (c >= 8 && c <= 16) cands.push_back({false, c, 64});
    } else if (context_table(*ctx, max_n, max_end, pc)) {
        cands.push_back({true, 0, 64, pc.table == ctx->small_table.get()});
    } else {
        cands.push_back({false, 0, 64});
    }
    // (the plan has its own engine: it only reads the context's points)
    if (cands.size() == 1) {
        plan->chosen = cands[0];
    } else {
        float best = 1e30f;
        for (const Candidate &c : cands) {
            const float ms = plan->time_msm(c, st);
            if (ms < best) {
                best = ms;
                plan->chosen = c;
            }
        }
        plan->tuned_ms = best;
    }
    plan->prepare_msm(plan->chosen, st);
    MH_HIP(hipStreamSynchronize(st));

    if (flags & kPlanGraph) {
        MH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            plan->ntt.enqueue(st);
            plan->engine.enqueue(st, nullptr);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        MH_HIP(hipStreamEndCapture(st, &plan->graph));
        MH_HIP(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
    }
    *out = plan.release();
    return ok();
    MH_CATCH
}

// One run: every instance's inverse NTT (in place in the device vector),
// then its MSM; out[i] is instance i's result (Jacobian).
extern "C" RustError midnight_hip_batch_plan_run(void *plan_ptr, RawG1Projective *out)
{
    if (!plan_ptr || !out) return make_error((int)hipErrorInvalidValue, "invalid batch plan run arguments");
    auto *plan = static_cast<BatchPlan *>(plan_ptr);
    MH_TRY
    std::lock_guard<std::mutex> g(plan->mu);
    hipStream_t st = lane_stream(plan->lane);
    ProfileScope prof;
    if (plan->exec) {
        if (profiling()) prof.begin("batch_graph", st);
        MH_HIP(hipGraphLaunch(plan->exec, st));
        if (profiling()) prof.end(st);
    } else {
        plan->ntt.enqueue(st);
        plan->engine.enqueue(st, profiling() ? &prof : nullptr);
    }
    MH_HIP(hipStreamSynchronize(st));
    prof.flush();
    plan->engine.finish(out);
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_hip_batch_plan_info(const void *plan_ptr, uint32_t *window_bits, uint32_t *chunk_len,
                                                  uint32_t *precomputed, uint32_t *graph, float *tuned_ms)
{
    if (!plan_ptr) return make_error((int)hipErrorInvalidValue, "null batch plan");
    const auto *plan = static_cast<const BatchPlan *>(plan_ptr);
    if (window_bits) *window_bits = plan->engine.last_window;
    if (chunk_len) *chunk_len = plan->chosen.chunk_len;
    if (precomputed) *precomputed = plan->chosen.precomputed ? 1 : 0;
    if (graph) *graph = plan->exec ? 1 : 0;
    if (tuned_ms) *tuned_ms = plan->tuned_ms;
    return ok();
}

extern "C" void midnight_hip_batch_plan_free(void *plan_ptr)
{
    auto *plan = static_cast<BatchPlan *>(plan_ptr);
    if (!plan) return;
    try {
        (void)hipStreamSynchronize(lane_stream(plan->lane));
    } catch (...) {
    }
    delete plan;
}
