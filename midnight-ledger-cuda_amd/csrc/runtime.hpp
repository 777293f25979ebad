// Host runtime shared by the backend's translation units: error type that
// becomes the reference's RustError at the C boundary, per-lane HIP
// streams, and RAII device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/midnight_hip.h"

namespace mh {

// A failed HIP call or a rejected argument.  `code` is the hipError_t value
// (or -1 for argument errors), as the reference returns cudaError_t codes.
struct hip_error : std::runtime_error {
    int code;
    hip_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define MH_HIP(expr)                                                              \
    do {                                                                          \
        hipError_t mh_e_ = (expr);                                                \
        if (mh_e_ != hipSuccess)                                                  \
            throw ::mh::hip_error((int)mh_e_, std::string(#expr) + ": " +         \
                                                  hipGetErrorString(mh_e_));      \
    } while (0)

#define MH_CHECK(cond, code, msg)                                                 \
    do {                                                                          \
        if (!(cond)) throw ::mh::hip_error((code), (msg));                        \
    } while (0)

RustError make_error(int code, const char *msg);
inline RustError ok() { return RustError{0, nullptr}; }

// Body of a C-ABI entry point: HIP failures and rejected arguments become
// their error code, host allocation failures hipErrorOutOfMemory.
#define MH_TRY try {
#define MH_CATCH                                                                            \
    }                                                                                       \
    catch (const ::mh::hip_error &e) { return ::mh::make_error(e.code, e.what()); }         \
    catch (const std::bad_alloc &) {                                                        \
        return ::mh::make_error((int)hipErrorOutOfMemory, "host allocation failed");        \
    }

// Proof execution lanes (the reference keeps one private device context per
// lane, bls12_381_g1_msm.cu:46-146).  Here a lane is one HIP stream; every
// producer and consumer of a lane's device vectors runs on that stream, so
// stream order is the only ordering the lane needs.
constexpr uint32_t kMaxLanes = 8;
// Lanes that exist internally: 0 .. kMaxLanes are the ones callers select
// (as in the reference); the proof server may run up to kServerSlots
// workers on lanes 1 .. kServerSlots, beyond the public range.
constexpr uint32_t kServerSlots = 16;
constexpr uint32_t kLaneCount = kServerSlots + 1;
uint32_t current_lane();
void set_current_lane(uint32_t lane);
hipStream_t lane_stream(uint32_t lane);
inline hipStream_t cur_stream() { return lane_stream(current_lane()); }
int device_id();

// Stream-ordered device memory: allocations come from the device's default
// memory pool on the current lane's stream and are returned to it on that
// stream.  hipFree would synchronise every stream of the device, so proofs
// running on different lanes would stall each other at every temporary
// buffer; the pool keeps freed memory for reuse (release threshold raised
// once per device).  In front of the pool, a per-stream cache hands a freed
// block to the next allocation of its size class on the same stream.
void *dev_alloc(size_t bytes, hipStream_t st);
void dev_free(void *p, hipStream_t st);

// Owning device allocation (used on the stream that allocated it).
template <class T>
struct DevBuf {
    T *ptr = nullptr;
    size_t len = 0;
    hipStream_t st = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), len(o.len), st(o.st) { o.ptr = nullptr; o.len = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            ptr = o.ptr;
            len = o.len;
            st = o.st;
            o.ptr = nullptr;
            o.len = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n)
    {
        release();
        if (n) {
            st = cur_stream();
            ptr = static_cast<T *>(dev_alloc(n * sizeof(T), st));
        }
        len = n;
    }
    // grow-only reuse for scratch buffers
    void ensure(size_t n)
    {
        if (n > len) alloc(n);
    }
    void release()
    {
        if (ptr) dev_free(ptr, st);
        ptr = nullptr;
        len = 0;
    }
    T *get() const { return ptr; }
};

inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

}  // namespace mh
