// Optional per-kernel timing with HIP events on the launching stream, for
// roofline reporting.  Disabled by default; when enabled, timed regions are
// bracketed by events and accumulated once the stream has synchronised.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace mh {

bool profiling();

// Collects (name, start, stop) event pairs; `flush` (after the stream has
// synchronised) adds their elapsed times to the registry.
class ProfileScope {
public:
    void begin(const char *name, hipStream_t st);
    void end(hipStream_t st);
    void flush();
    ~ProfileScope();

private:
    struct Rec {
        const char *name;
        hipEvent_t a, b;
        bool open;
    };
    std::vector<Rec> recs_;
};

}  // namespace mh
