// Proof server: a request loop that keeps keys and the SRS resident and
// proves concurrent requests of different sizes on one GPU.
//
// Each slot is a worker thread bound to its own lane (HIP stream), so the
// stages of proofs admitted together overlap on the device (MSM contexts
// keep one engine per lane, every scratch buffer is per lane, temporary
// device memory is stream-ordered).  Admission is continuous: a queued
// request starts as soon as a slot is free and its HBM footprint fits in the
// budget beside the proofs already running (first fit in arrival order; a
// request passed over too often blocks later ones until it fits; one
// request always runs when nothing else does).  Results do not depend on
// what runs beside a proof: the arithmetic is exact and every proof's
// randomness comes from its own seed.
//
// The reference serves proofs from proof-server (src/worker_pool.rs) with a
// pool of CPU workers each calling the prover; this is the GPU-side
// equivalent of that pool for the proof pipeline in prover.hip.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "runtime.hpp"

namespace mh {

struct PlonkKey;
std::vector<uint8_t> plonk_prove(PlonkKey &pk, const RawFr *advice, const RawFr *instance,
                                 const size_t *instance_lens, uint64_t seed);
size_t plonk_footprint(const PlonkKey &pk);
size_t plonk_advice_len(const PlonkKey &pk);
size_t plonk_instance_cols(const PlonkKey &pk);

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct ProofJob {
    uint64_t id = 0;
    PlonkKey *key = nullptr;
    std::vector<RawFr> advice, instance;
    std::vector<size_t> instance_lens;   // one per instance column of the key
    uint64_t seed = 0;
    size_t footprint = 0;
    double t_submit = 0, t_start = 0, t_end = 0;
    uint32_t skipped = 0;
    int state = 0;   // 0 queued, 1 running, 2 done, 3 failed
    std::vector<uint8_t> proof;
    int err_code = 0;
    std::string err;
};

}  // namespace

struct ProofServer {
    uint32_t slots = 1;
    size_t budget = 0, in_use = 0;
    uint32_t running = 0;
    bool stopping = false;
    uint64_t next_id = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<ProofJob>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<ProofJob>> jobs;
    std::vector<std::thread> workers;

    // first queued job that fits; the head blocks the others once it has
    // been passed over 2 * slots times
    std::shared_ptr<ProofJob> take_fitting()
    {
        for (size_t i = 0; i < queue.size(); ++i) {
            auto &j = queue[i];
            const bool fits = running == 0 || in_use + j->footprint <= budget;
            if (fits) {
                auto job = j;
                for (size_t h = 0; h < i; ++h) queue[h]->skipped++;
                queue.erase(queue.begin() + (long)i);
                return job;
            }
            if (i == 0 && j->skipped >= 2 * slots) break;
        }
        return nullptr;
    }

    void worker(uint32_t lane)
    {
        set_current_lane(lane);
        for (;;) {
            std::shared_ptr<ProofJob> job;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&]() { return stopping || (!queue.empty() && (job = take_fitting()) != nullptr); });
                if (!job) return;   // stopping with nothing admitted
                in_use += job->footprint;
                running++;
                job->state = 1;
                job->t_start = now_ms();
            }
            std::vector<uint8_t> proof;
            int state = 2, code = 0;
            std::string err;
            try {
                proof = plonk_prove(*job->key, job->advice.data(), job->instance.data(), job->instance_lens.data(),
                                    job->seed);
            } catch (const hip_error &e) {
                code = e.code;
                err = e.what();
                state = 3;
            } catch (const std::exception &e) {
                code = -1;
                err = e.what();
                state = 3;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                job->proof = std::move(proof);
                job->err_code = code;
                job->err = err;
                job->state = state;
                job->t_end = now_ms();
                job->advice.clear();
                job->advice.shrink_to_fit();
                in_use -= job->footprint;
                running--;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

}  // namespace mh

using namespace mh;

extern "C" RustError midnight_hip_proof_server_create(void **out, uint32_t slots, uint64_t hbm_budget_bytes)
{
    if (!out || slots == 0 || slots > kServerSlots)
        return make_error((int)hipErrorInvalidValue, "slots must be between 1 and 16");
    *out = nullptr;
    try {
        auto s = std::make_unique<ProofServer>();
        s->slots = slots;
        s->budget = hbm_budget_bytes;
        for (uint32_t i = 0; i < slots; ++i) s->workers.emplace_back([p = s.get(), i]() { p->worker(i + 1); });
        *out = s.release();
        return ok();
    } catch (const std::exception &e) {
        return make_error(-1, e.what());
    }
}

extern "C" RustError midnight_hip_proof_server_submit_with_instances(void *server, void *pk, const RawFr *advice,
                                                                     size_t advice_len, const RawFr *instance,
                                                                     const size_t *instance_lens,
                                                                     size_t instance_cols, uint64_t seed,
                                                                     uint64_t *job_id)
{
    if (!server || !pk || !advice || !job_id) return make_error((int)hipErrorInvalidValue, "invalid submit arguments");
    auto *s = reinterpret_cast<ProofServer *>(server);
    auto *key = reinterpret_cast<PlonkKey *>(pk);
    if (advice_len != plonk_advice_len(*key))
        return make_error((int)hipErrorInvalidValue, "advice length does not match the key");
    if (instance_cols != plonk_instance_cols(*key) || (instance_cols && !instance_lens))
        return make_error((int)hipErrorInvalidValue, "instance columns do not match the key");
    size_t total = 0;
    for (size_t c = 0; c < instance_cols; ++c) total += instance_lens[c];
    if (total && !instance) return make_error((int)hipErrorInvalidValue, "null instance values");
    try {
        auto job = std::make_shared<ProofJob>();
        job->key = key;
        job->advice.assign(advice, advice + advice_len);
        if (total) job->instance.assign(instance, instance + total);
        job->instance_lens.assign(instance_lens, instance_lens + instance_cols);
        job->seed = seed;
        job->footprint = plonk_footprint(*key);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            job->id = s->next_id++;
            job->t_submit = now_ms();
            s->queue.push_back(job);
            s->jobs[job->id] = job;
        }
        s->cv_work.notify_all();
        *job_id = job->id;
        return ok();
    } catch (const std::exception &e) {
        return make_error(-1, e.what());
    }
}

extern "C" RustError midnight_hip_proof_server_submit(void *server, void *pk, const RawFr *advice, size_t advice_len,
                                                      uint64_t seed, uint64_t *job_id)
{
    if (!pk) return make_error((int)hipErrorInvalidValue, "invalid submit arguments");
    const size_t I = plonk_instance_cols(*reinterpret_cast<PlonkKey *>(pk));
    std::vector<size_t> empty(I, 0);
    return midnight_hip_proof_server_submit_with_instances(server, pk, advice, advice_len, nullptr,
                                                          I ? empty.data() : nullptr, I, seed, job_id);
}

// Blocks until the job is finished; copies the proof and its queueing and
// proving times, and forgets the job.  A null or too small proof buffer
// returns an error with *proof_len set and keeps the finished job, so the
// caller can wait on it again with a buffer of that size.
extern "C" RustError midnight_hip_proof_server_wait(void *server, uint64_t job_id, uint8_t *proof, size_t capacity,
                                                    size_t *proof_len, double *queue_ms, double *prove_ms)
{
    if (!server || !proof_len) return make_error((int)hipErrorInvalidValue, "invalid wait arguments");
    auto *s = reinterpret_cast<ProofServer *>(server);
    std::shared_ptr<ProofJob> job;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(job_id);
        if (it == s->jobs.end()) return make_error((int)hipErrorInvalidValue, "unknown job");
        job = it->second;
        s->cv_done.wait(lk, [&]() { return job->state >= 2; });
        if (queue_ms) *queue_ms = job->t_start - job->t_submit;
        if (prove_ms) *prove_ms = job->t_end - job->t_start;
        if (job->state == 2) {
            *proof_len = job->proof.size();
            if (!proof || capacity < job->proof.size())
                return make_error((int)hipErrorInvalidValue, "proof buffer too small");
        }
        s->jobs.erase(job_id);
    }
    if (job->state == 3) return make_error(job->err_code, job->err.c_str());
    std::memcpy(proof, job->proof.data(), job->proof.size());
    return ok();
}

// Finishes the queued and running proofs, then stops the workers.
extern "C" void midnight_hip_proof_server_free(void *server)
{
    auto *s = reinterpret_cast<ProofServer *>(server);
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&]() { return s->queue.empty() && s->running == 0; });
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    delete s;
}
