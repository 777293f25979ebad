// Per-kernel timing registry (see profile.hpp).
#include <atomic>
#include <map>
#include <mutex>
#include <string>

#include "profile.hpp"
#include "runtime.hpp"

namespace mh {

namespace {
std::atomic<bool> g_on{false};
std::mutex g_mu;
std::map<std::string, std::pair<double, uint64_t>> g_reg;
}  // namespace

bool profiling() { return g_on.load(std::memory_order_relaxed); }

void ProfileScope::begin(const char *name, hipStream_t st)
{
    if (!profiling()) return;
    Rec r{name, nullptr, nullptr, true};
    MH_HIP(hipEventCreate(&r.a));
    MH_HIP(hipEventCreate(&r.b));
    MH_HIP(hipEventRecord(r.a, st));
    recs_.push_back(r);
}

void ProfileScope::end(hipStream_t st)
{
    if (!profiling()) return;
    for (size_t i = recs_.size(); i-- > 0;) {
        if (recs_[i].open) {
            MH_HIP(hipEventRecord(recs_[i].b, st));
            recs_[i].open = false;
            return;
        }
    }
}

void ProfileScope::flush()
{
    std::lock_guard<std::mutex> g(g_mu);
    for (auto &r : recs_) {
        float ms = 0;
        if (!r.open && hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto &e = g_reg[r.name];
            e.first += ms;
            e.second += 1;
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    recs_.clear();
}

ProfileScope::~ProfileScope()
{
    for (auto &r : recs_) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
}

}  // namespace mh

extern "C" void midnight_hip_profile_enable(int on)
{
    std::lock_guard<std::mutex> g(mh::g_mu);
    mh::g_reg.clear();
    mh::g_on.store(on != 0);
}

extern "C" int midnight_hip_profile_read(const char *name, double *total_ms, uint64_t *count)
{
    std::lock_guard<std::mutex> g(mh::g_mu);
    auto it = mh::g_reg.find(name ? name : "");
    if (it == mh::g_reg.end()) {
        if (total_ms) *total_ms = 0;
        if (count) *count = 0;
        return 0;
    }
    if (total_ms) *total_ms = it->second.first;
    if (count) *count = it->second.second;
    return 1;
}
