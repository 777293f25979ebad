// C ABI: lanes, device vectors, NTT / coset and Fr vector entry points.
// Argument checks, return codes and messages follow the reference entry
// points cited at each function (midnight-sppark-msm/src/bls12_381_g1_msm.cu).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "devvec.hpp"
#include "frops.hpp"
#include "ntt.hpp"
#include "runtime.hpp"

namespace mh {

RustError make_error(int code, const char *msg)
{
    RustError e;
    e.code = code == 0 ? -1 : code;
    e.message = msg ? strdup(msg) : nullptr;
    return e;
}

namespace {
thread_local uint32_t t_lane = 0;
hipStream_t g_streams[64][kLaneCount];
std::mutex g_stream_mu;
}  // namespace

uint32_t current_lane() { return t_lane; }
void set_current_lane(uint32_t lane) { t_lane = lane; }

int device_id()
{
    int d = 0;
    MH_HIP(hipGetDevice(&d));
    return d;
}

// Streams are created once per (device, lane) and live until the process
// ends, so each thread caches the ones it has seen: the shared table and
// its mutex are touched once per thread and lane, not on every call.
hipStream_t lane_stream(uint32_t lane)
{
    MH_CHECK(lane < kLaneCount, (int)hipErrorInvalidValue, "invalid execution lane");
    const int dev = device_id();
    thread_local int t_dev = -1;
    thread_local hipStream_t t_cache[kLaneCount] = {};
    if (t_dev != dev) {
        for (auto &x : t_cache) x = nullptr;
        t_dev = dev;
    }
    if (t_cache[lane]) return t_cache[lane];
    std::lock_guard<std::mutex> g(g_stream_mu);
    hipStream_t &s = g_streams[dev][lane];
    if (!s) MH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    t_cache[lane] = s;
    return s;
}

namespace {

// Per-stream caches of freed blocks.  A block freed on a stream is handed
// to the next allocation of its size class on the same stream: stream order
// puts every use of the new owner after every use of the old one, as the
// pool's own reuse would, without the pool's device-wide lock on every call
// (a proof makes about 70 temporary allocations; twelve lanes proving at
// once spent longer queued on that lock than on their kernels).  Blocks up
// to kCacheBlockMax are kept, up to kCacheStreamMax bytes per stream; the
// rest go back to the pool.  Cached memory stays with the process.
constexpr size_t kCacheBlockMax = size_t(1) << 30;
constexpr size_t kCacheStreamMax = size_t(8) << 30;

// size classes in quarter steps of powers of two (at most 25% over)
size_t size_class(size_t bytes)
{
    if (bytes <= 256) return 256;
    int lg = 63 - __builtin_clzll(bytes - 1);   // 2^lg < bytes <= 2^(lg+1)
    const size_t step = (size_t(1) << lg) >> 2;
    return ((bytes + step - 1) / step) * step;
}

struct StreamCache {
    std::unordered_multimap<size_t, void *> free;
    size_t bytes = 0;
};

std::mutex g_cache_mu;
std::unordered_map<hipStream_t, StreamCache> g_cache;
std::unordered_map<void *, size_t> g_block_class;   // every live or cached block

void pool_setup()
{
    static std::once_flag pool_once[64];
    const int dev = device_id();
    std::call_once(pool_once[dev], [dev]() {
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t keep = ~0ull;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    });
}

}  // namespace

void *dev_alloc(size_t bytes, hipStream_t st)
{
    const size_t cls = size_class(bytes);
    {
        std::lock_guard<std::mutex> g(g_cache_mu);
        auto c = g_cache.find(st);
        if (c != g_cache.end()) {
            auto it = c->second.free.find(cls);
            if (it != c->second.free.end()) {
                void *p = it->second;
                c->second.free.erase(it);
                c->second.bytes -= cls;
                return p;
            }
        }
    }
    pool_setup();
    void *p = nullptr;
    if (hipMallocAsync(&p, cls, st) != hipSuccess) {
        // out of memory: hand every cached block back to the pool, wait for
        // the frees to complete, and try once more
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> g(g_cache_mu);
            for (auto &c : g_cache) {
                for (auto &blk : c.second.free) {
                    g_block_class.erase(blk.second);
                    (void)hipFreeAsync(blk.second, c.first);
                }
                c.second.free.clear();
                c.second.bytes = 0;
            }
        }
        MH_HIP(hipDeviceSynchronize());
        MH_HIP(hipMallocAsync(&p, cls, st));
    }
    std::lock_guard<std::mutex> g(g_cache_mu);
    g_block_class[p] = cls;
    return p;
}

void dev_free(void *p, hipStream_t st)
{
    {
        std::lock_guard<std::mutex> g(g_cache_mu);
        auto b = g_block_class.find(p);
        if (b != g_block_class.end()) {
            const size_t cls = b->second;
            StreamCache &c = g_cache[st];
            if (cls <= kCacheBlockMax && c.bytes + cls <= kCacheStreamMax) {
                c.free.emplace(cls, p);
                c.bytes += cls;
                return;
            }
            g_block_class.erase(b);
        }
    }
    (void)hipFreeAsync(p, st);
}

// per-lane reusable staging/scratch buffers for the host-bridge entry points
namespace {
struct LaneScratch {
    std::mutex mu;
    DevBuf<Fr> a, b, c;
};
LaneScratch g_scratch[64][kLaneCount];
}  // namespace

LaneScratch &lane_scratch() { return g_scratch[device_id()][current_lane()]; }

Fr to_fr(const RawFr &r)
{
    Fr f;
    std::memcpy(f.v, r.limbs, 32);
    return f;
}

HostFr to_hfr(const RawFr &r)
{
    HostFr h;
    std::memcpy(h.l, r.limbs, 32);
    return h;
}

}  // namespace mh

using namespace mh;

extern "C" RustError midnight_sppark_execution_lane_set(uint32_t lane)
{
    if (lane > kMaxLanes) return make_error((int)hipErrorInvalidValue, "invalid sppark proof execution lane");
    MH_TRY
    set_current_lane(lane);
    (void)lane_stream(lane);
    return ok();
    MH_CATCH
}

// ---------------------------------------------------------------- device vectors

extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_upload(void **out, const RawFr *values, size_t len)
{
    if (!out) return make_error((int)hipErrorInvalidValue, "out is null");
    *out = nullptr;
    if (len && !values) return make_error((int)hipErrorInvalidValue, "values is null");
    MH_TRY
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = len;
    dv->values.alloc(len);
    if (len) {
        hipStream_t st = cur_stream();
        MH_HIP(hipMemcpyAsync(dv->values.get(), values, len * 32, hipMemcpyHostToDevice, st));
        MH_HIP(hipStreamSynchronize(st));
    }
    *out = dv.release();
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_upload_many(void **out, const RawFr *const *values,
                                                                         const size_t *lengths, size_t count)
{
    if (!out || (count && (!values || !lengths)))
        return make_error((int)hipErrorInvalidValue, "invalid multi-slice upload arguments");
    *out = nullptr;
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) {
        if ((lengths[i] && !values[i]) || lengths[i] > SIZE_MAX - total)
            return make_error((int)hipErrorInvalidValue, "invalid multi-slice upload source");
        total += lengths[i];
    }
    MH_TRY
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = total;
    dv->values.alloc(total);
    hipStream_t st = cur_stream();
    size_t off = 0;
    for (size_t i = 0; i < count; ++i) {
        if (!lengths[i]) continue;
        MH_HIP(hipMemcpyAsync(dv->values.get() + off, values[i], lengths[i] * 32, hipMemcpyHostToDevice, st));
        off += lengths[i];
    }
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_download(RawFr *out, const void *device_vec, size_t len)
{
    if (!device_vec || (len && !out))
        return make_error((int)hipErrorInvalidValue, "invalid device vector download arguments");
    const DeviceVec *dv = static_cast<const DeviceVec *>(device_vec);
    if (dv->len != len) return make_error((int)hipErrorInvalidValue, "device vector length mismatch");
    MH_TRY
    hipStream_t st = cur_stream();
    if (len) MH_HIP(hipMemcpyAsync(out, dv->values.get(), len * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

extern "C" void midnight_sppark_bls12_381_fr_device_vec_free(void *device_vec)
{
    DeviceVec *dv = static_cast<DeviceVec *>(device_vec);
    if (!dv) return;
    // pending work of the owning lane may still read the buffer
    try {
        (void)hipStreamSynchronize(lane_stream(dv->lane));
    } catch (...) {
    }
    delete dv;
}

namespace {

__global__ void k_scatter(Fr *out, const uint32_t *idx, const Fr *vals, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) store_fr(out + idx[i], load_fr(vals + i));
}

// per polynomial (block): index of the first nonzero value (len if none) and
// the number of nonzero values
__global__ __launch_bounds__(256) void k_nonzero_stats(uint64_t *starts, uint64_t *counts, const Fr *v, size_t len)
{
    __shared__ unsigned long long first;
    __shared__ unsigned long long live;
    if (threadIdx.x == 0) {
        first = len;
        live = 0;
    }
    __syncthreads();
    const Fr *p = v + (size_t)blockIdx.x * len;
    unsigned long long f = len, c = 0;
    for (size_t i = threadIdx.x; i < len; i += blockDim.x)
        if (!load_fr(p + i).is_zero()) {
            if (i < f) f = i;
            ++c;
        }
    atomicMin(&first, f);
    atomicAdd(&live, c);
    __syncthreads();
    if (threadIdx.x == 0) {
        starts[blockIdx.x] = first;
        counts[blockIdx.x] = live;
    }
}

std::unique_ptr<DeviceVec> alloc_vec(size_t len)
{
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = len;
    dv->values.alloc(std::max<size_t>(len, 1));
    return dv;
}

}  // namespace

// bls12_381_g1_msm.cu:2878-2914
extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_scatter_host(void *device_vec, const uint32_t *indices,
                                                                          const RawFr *values, size_t count)
{
    if (device_vec == nullptr || (count != 0 && (indices == nullptr || values == nullptr)))
        return make_error((int)hipErrorInvalidValue, "invalid device scatter arguments");
    auto *dv = static_cast<DeviceVec *>(device_vec);
    for (size_t i = 0; i < count; ++i)
        if (indices[i] >= dv->len) return make_error((int)hipErrorInvalidValue, "device scatter index out of range");
    if (count == 0) return ok();
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<uint32_t> d_idx(count);
    DevBuf<Fr> d_vals(count);
    MH_HIP(hipMemcpyAsync(d_idx.get(), indices, count * 4, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(d_vals.get(), values, count * 32, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)ceil_div(count, 256)), dim3(256), 0, st, dv->values.get(),
                       (const uint32_t *)d_idx.get(), (const Fr *)d_vals.get(), count);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:2916-2962
extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_concat_host(void **out, const void *prefix_device_vec,
                                                                         const RawFr *suffix, size_t suffix_len)
{
    if (out == nullptr || prefix_device_vec == nullptr || (suffix_len != 0 && suffix == nullptr))
        return make_error((int)hipErrorInvalidValue, "invalid device vector concat arguments");
    *out = nullptr;
    const auto *prefix = static_cast<const DeviceVec *>(prefix_device_vec);
    if (suffix_len > SIZE_MAX - prefix->len)
        return make_error((int)hipErrorInvalidValue, "device vector concat length overflow");
    MH_TRY
    hipStream_t st = cur_stream();
    auto dv = alloc_vec(prefix->len + suffix_len);
    if (prefix->len)
        MH_HIP(hipMemcpyAsync(dv->values.get(), prefix->values.get(), prefix->len * 32, hipMemcpyDeviceToDevice, st));
    if (suffix_len)
        MH_HIP(hipMemcpyAsync(dv->values.get() + prefix->len, suffix, suffix_len * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:2964-3023
extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_concat_many_host(void **out,
                                                                              const void *const *source_device_vecs,
                                                                              size_t source_count, const RawFr *suffix,
                                                                              size_t suffix_len)
{
    if (out == nullptr || source_count == 0 || source_device_vecs == nullptr || (suffix_len != 0 && suffix == nullptr))
        return make_error((int)hipErrorInvalidValue, "invalid multi-device vector concat arguments");
    *out = nullptr;
    size_t total = suffix_len;
    for (size_t i = 0; i < source_count; ++i) {
        if (source_device_vecs[i] == nullptr)
            return make_error((int)hipErrorInvalidValue, "null device vector concat source");
        const auto *src = static_cast<const DeviceVec *>(source_device_vecs[i]);
        if (src->len > SIZE_MAX - total)
            return make_error((int)hipErrorInvalidValue, "multi-device vector concat length overflow");
        total += src->len;
    }
    MH_TRY
    hipStream_t st = cur_stream();
    auto dv = alloc_vec(total);
    size_t off = 0;
    for (size_t i = 0; i < source_count; ++i) {
        const auto *src = static_cast<const DeviceVec *>(source_device_vecs[i]);
        if (src->len)
            MH_HIP(hipMemcpyAsync(dv->values.get() + off, src->values.get(), src->len * 32, hipMemcpyDeviceToDevice,
                                  st));
        off += src->len;
    }
    if (suffix_len)
        MH_HIP(hipMemcpyAsync(dv->values.get() + off, suffix, suffix_len * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:3025-3072
extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_gather_polys(void **out, const void *source_device_vec,
                                                                          size_t poly_len, const uint32_t *indices,
                                                                          size_t index_count)
{
    if (out == nullptr || source_device_vec == nullptr || poly_len == 0 || indices == nullptr || index_count == 0 ||
        index_count > SIZE_MAX / poly_len)
        return make_error((int)hipErrorInvalidValue, "invalid device polynomial gather arguments");
    *out = nullptr;
    const auto *src = static_cast<const DeviceVec *>(source_device_vec);
    if (src->len % poly_len != 0)
        return make_error((int)hipErrorInvalidValue, "device polynomial gather source shape mismatch");
    const size_t source_count = src->len / poly_len;
    for (size_t i = 0; i < index_count; ++i)
        if (indices[i] >= source_count)
            return make_error((int)hipErrorInvalidValue, "device polynomial gather index out of range");
    MH_TRY
    hipStream_t st = cur_stream();
    auto dv = alloc_vec(index_count * poly_len);
    for (size_t i = 0; i < index_count; ++i)
        MH_HIP(hipMemcpyAsync(dv->values.get() + i * poly_len, src->values.get() + (size_t)indices[i] * poly_len,
                              poly_len * 32, hipMemcpyDeviceToDevice, st));
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:3101-3179
extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_download_nonzero_suffixes(
    RawFr *out, size_t out_capacity, size_t *out_len, size_t *starts, size_t *nonzero_counts, const void *device_vec,
    size_t scalar_offset, size_t poly_len, size_t count)
{
    if (out_len == nullptr || starts == nullptr || device_vec == nullptr || poly_len == 0 || count == 0 ||
        count > SIZE_MAX / poly_len || (out_capacity != 0 && out == nullptr))
        return make_error((int)hipErrorInvalidValue, "invalid nonzero suffix download arguments");
    *out_len = 0;
    const auto *dv = static_cast<const DeviceVec *>(device_vec);
    const size_t required = poly_len * count;
    if (scalar_offset > dv->len || required > dv->len - scalar_offset)
        return make_error((int)hipErrorInvalidValue, "nonzero suffix range exceeds device vector");
    if (count > 0x7fffffffu) return make_error((int)hipErrorInvalidValue, "too many polynomials");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<uint64_t> d_stats(2 * count);
    const Fr *base = dv->values.get() + scalar_offset;
    hipLaunchKernelGGL(k_nonzero_stats, dim3((unsigned)count), dim3(256), 0, st, d_stats.get(),
                       d_stats.get() + count, base, poly_len);
    MH_HIP(hipGetLastError());
    std::vector<uint64_t> h(2 * count);
    MH_HIP(hipMemcpyAsync(h.data(), d_stats.get(), 2 * count * 8, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) {
        starts[i] = (size_t)h[i];
        if (nonzero_counts) nonzero_counts[i] = (size_t)h[count + i];
        total += poly_len - starts[i];
    }
    if (out == nullptr && out_capacity == 0) {
        *out_len = total;
        return ok();
    }
    if (total > out_capacity) return make_error((int)hipErrorInvalidValue, "nonzero suffix output is too small");
    size_t cursor = 0;
    for (size_t i = 0; i < count; ++i) {
        const size_t n = poly_len - starts[i];
        if (n)
            MH_HIP(hipMemcpyAsync(out + cursor, base + i * poly_len + starts[i], n * 32, hipMemcpyDeviceToHost, st));
        cursor += n;
    }
    MH_HIP(hipStreamSynchronize(st));
    *out_len = total;
    return ok();
    MH_CATCH
}

extern "C" size_t midnight_hip_fr_device_vec_len(const void *device_vec)
{
    return device_vec ? static_cast<const DeviceVec *>(device_vec)->len : 0;
}

extern "C" RustError midnight_sppark_bls12_381_fr_device_quotient_pieces(void **out, const void *quotient_vec,
                                                                         size_t poly_len, size_t piece_count,
                                                                         const RawFr *blinds, size_t blind_count)
{
    if (!out || !quotient_vec || poly_len < 2 || piece_count < 2 || blind_count != piece_count - 1 || !blinds ||
        piece_count > SIZE_MAX / poly_len)
        return make_error((int)hipErrorInvalidValue, "invalid resident quotient piece arguments");
    *out = nullptr;
    const DeviceVec *q = static_cast<const DeviceVec *>(quotient_vec);
    if (q->len < (poly_len - 1) * piece_count)
        return make_error((int)hipErrorInvalidValue, "resident quotient is too short");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_blinds(blind_count);
    MH_HIP(hipMemcpyAsync(d_blinds.get(), blinds, blind_count * 32, hipMemcpyHostToDevice, st));
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = poly_len * piece_count;
    dv->values.alloc(dv->len);
    fr_quotient_pieces(dv->values.get(), q->values.get(), poly_len, piece_count, d_blinds.get(), st);
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}

// ---------------------------------------------------------------- NTT family

namespace {

// Host-bridge NTT: upload, transform in place on the lane's stream, download.
RustError host_ntt(RawFr *values, uint32_t log_n, NttJob job, const RawFr *factors, size_t factor_len)
{
    const size_t n = size_t(1) << log_n;
    LaneScratch &ls = lane_scratch();
    std::lock_guard<std::mutex> g(ls.mu);
    hipStream_t st = cur_stream();
    ls.a.ensure(n);
    ls.b.ensure(ntt_scratch_elems(log_n, 1));
    MH_HIP(hipMemcpyAsync(ls.a.get(), values, n * 32, hipMemcpyHostToDevice, st));
    DevBuf<Fr> d_factors;
    if (factors) {
        d_factors.alloc(factor_len);
        MH_HIP(hipMemcpyAsync(d_factors.get(), factors, factor_len * 32, hipMemcpyHostToDevice, st));
        job.factors = d_factors.get();
        job.factor_len = factor_len;
        job.pre = 2;
    }
    job.in = ls.a.get();
    job.out = ls.a.get();
    job.scratch = ls.b.get();
    job.in_pstride = job.out_pstride = n;
    job.log_n = log_n;
    job.count = 1;
    ntt_run(job, st);
    MH_HIP(hipMemcpyAsync(values, ls.a.get(), n * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
}

}  // namespace

extern "C" RustError midnight_sppark_bls12_381_fr_ntt(RawFr *values, uint32_t log_n, bool inverse)
{
    if (!values && log_n) return make_error((int)hipErrorInvalidValue, "values is null");
    MH_TRY
    NttJob job;
    job.inverse = inverse;
    return host_ntt(values, log_n, job, nullptr, 0);
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_ntt_zeta(RawFr *values, uint32_t log_n, bool inverse, RawFr scale,
                                                           RawFr first, RawFr second)
{
    if (log_n == 0) return ok();
    MH_TRY
    NttJob job;
    job.inverse = inverse;
    job.post_zeta = true;
    job.scale = to_hfr(scale);
    job.first = to_hfr(first);
    job.second = to_hfr(second);
    return host_ntt(values, log_n, job, nullptr, 0);
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_divide_ntt_zeta(RawFr *values, uint32_t log_n, const RawFr *factors,
                                                                  size_t factor_len, RawFr scale, RawFr first,
                                                                  RawFr second)
{
    if (log_n == 0 || factor_len == 0) return ok();
    MH_TRY
    NttJob job;
    job.inverse = true;
    job.post_zeta = true;
    job.scale = to_hfr(scale);
    job.first = to_hfr(first);
    job.second = to_hfr(second);
    return host_ntt(values, log_n, job, factors, factor_len);
    MH_CATCH
}

namespace {

// coeff_to_extended into a fresh device buffer (count polys of 2^ext_log_n)
void coeff_to_extended_device(Fr *dst, const Fr *src, size_t input_len, size_t count, uint32_t ext_log_n,
                              const RawFr &first, const RawFr &second, hipStream_t st)
{
    LaneScratch &ls = lane_scratch();
    std::lock_guard<std::mutex> g(ls.mu);
    ls.c.ensure(ntt_scratch_elems(ext_log_n, count));
    NttJob job;
    job.in = src;
    job.out = dst;
    job.scratch = ls.c.get();
    job.in_pstride = input_len;
    job.out_pstride = size_t(1) << ext_log_n;
    job.log_n = ext_log_n;
    job.count = (uint32_t)count;
    job.pre = 1;
    job.in_len = input_len;
    job.first = to_hfr(first);
    job.second = to_hfr(second);
    ntt_run(job, st);
    MH_HIP(hipStreamSynchronize(st));
}

}  // namespace

extern "C" RustError midnight_sppark_bls12_381_fr_coeff_to_extended_batch(RawFr *out, const RawFr *inputs,
                                                                          size_t input_len, size_t count,
                                                                          uint32_t extended_log_n, RawFr first,
                                                                          RawFr second)
{
    if (input_len == 0 || count == 0) return ok();
    size_t output_len = size_t(1) << extended_log_n;
    if (input_len > output_len) return make_error(-1, "input length exceeds extended domain length");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_in(input_len * count), d_out(output_len * count);
    MH_HIP(hipMemcpyAsync(d_in.get(), inputs, input_len * count * 32, hipMemcpyHostToDevice, st));
    coeff_to_extended_device(d_out.get(), d_in.get(), input_len, count, extended_log_n, first, second, st);
    MH_HIP(hipMemcpyAsync(out, d_out.get(), output_len * count * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_device_coeff_to_extended_batch(void **out, const RawFr *inputs,
                                                                                 size_t input_len, size_t count,
                                                                                 uint32_t extended_log_n,
                                                                                 RawFr first, RawFr second)
{
    if (!out) return make_error((int)hipErrorInvalidValue, "out is null");
    *out = nullptr;
    if (input_len == 0 || count == 0 || !inputs)
        return make_error((int)hipErrorInvalidValue, "invalid coefficient batch");
    size_t output_len = size_t(1) << extended_log_n;
    if (input_len > output_len)
        return make_error((int)hipErrorInvalidValue, "input length exceeds extended domain length");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_in(input_len * count);
    MH_HIP(hipMemcpyAsync(d_in.get(), inputs, input_len * count * 32, hipMemcpyHostToDevice, st));
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = output_len * count;
    dv->values.alloc(dv->len);
    coeff_to_extended_device(dv->values.get(), d_in.get(), input_len, count, extended_log_n, first, second, st);
    // the handle may be consumed on another lane: it is returned complete
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_device_coeff_to_extended_from_device_batch(
    void **out, const void *source, size_t input_len, size_t count, uint32_t extended_log_n, RawFr first,
    RawFr second)
{
    if (!out || !source) return make_error((int)hipErrorInvalidValue, "invalid resident coefficient source");
    *out = nullptr;
    const DeviceVec *src = static_cast<const DeviceVec *>(source);
    if (input_len == 0 || count == 0 || src->len != input_len * count)
        return make_error((int)hipErrorInvalidValue, "resident coefficient source length mismatch");
    size_t output_len = size_t(1) << extended_log_n;
    if (input_len > output_len || count > SIZE_MAX / output_len)
        return make_error((int)hipErrorInvalidValue, "invalid resident extended-domain size");
    MH_TRY
    hipStream_t st = cur_stream();
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = output_len * count;
    dv->values.alloc(dv->len);
    coeff_to_extended_device(dv->values.get(), src->values.get(), input_len, count, extended_log_n, first, second,
                             st);
    // the handle may be consumed on another lane: it is returned complete
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}

namespace {

void device_ntt_zeta(DeviceVec *dv, uint32_t log_n, size_t count, bool inverse, const RawFr &scale,
                     const RawFr &first, const RawFr &second, const Fr *factors, size_t factor_len)
{
    hipStream_t st = cur_stream();
    const size_t n = size_t(1) << log_n;
    // reusable per-lane scratch: no allocation per transform
    LaneScratch &ls = lane_scratch();
    std::lock_guard<std::mutex> g(ls.mu);
    ls.c.ensure(ntt_scratch_elems(log_n, count));
    NttJob job;
    job.in = dv->values.get();
    job.out = dv->values.get();
    job.scratch = ls.c.get();
    job.in_pstride = job.out_pstride = n;
    job.log_n = log_n;
    job.count = (uint32_t)count;
    job.inverse = inverse;
    job.post_zeta = true;
    job.scale = to_hfr(scale);
    job.first = to_hfr(first);
    job.second = to_hfr(second);
    if (factors) {
        job.pre = 2;
        job.factors = factors;
        job.factor_len = factor_len;
    }
    ntt_run(job, st);
    // callers of the resident entry points see completed transforms
    MH_HIP(hipStreamSynchronize(st));
}

}  // namespace

extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_ntt_zeta(void *device_vec, uint32_t log_n, bool inverse,
                                                                      RawFr scale, RawFr first, RawFr second)
{
    if (!device_vec) return make_error((int)hipErrorInvalidValue, "device vector is null");
    DeviceVec *dv = static_cast<DeviceVec *>(device_vec);
    if (log_n >= 63 || dv->len != (size_t(1) << log_n))
        return make_error((int)hipErrorInvalidValue, "device vector length mismatch");
    if (log_n == 0) return ok();
    MH_TRY
    device_ntt_zeta(dv, log_n, 1, inverse, scale, first, second, nullptr, 0);
    return ok();
    MH_CATCH
}

// --------------------------------------------------- no-allocation variant
//
// The four-step transform stages the 2^log_n values between its two passes
// in a scratch buffer.  For callers that cannot allocate device memory, the
// variant below takes that buffer from this static array instead of the
// lane's grow-only scratch: 2^20 Fr (32 MiB) per GPU, held until the process
// ends.  Every call on one GPU shares it, so calls that overlap on one GPU
// (different lanes or threads at once) race; calls on one lane are ordered.
// The pass kernels overwrite every scratch element before reading it, so
// nothing needs clearing between calls.
constexpr uint32_t kNoAllocMaxLog = 20;
static __device__ __align__(256) uint32_t scratch[(size_t(1) << kNoAllocMaxLog) * 8];

namespace {

// device address of `scratch` on the current device, looked up once per device
Fr *noalloc_scratch()
{
    static std::once_flag once[64];
    static Fr *ptr[64];
    const int dev = device_id();
    std::call_once(once[dev], [dev]() {
        void *p = nullptr;
        MH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = static_cast<Fr *>(p);
    });
    return ptr[dev];
}

}  // namespace

// midnight_sppark_bls12_381_fr_device_vec_ntt_zeta with no allocation: the
// same transform of the same vector (limb-identical), its staging in the
// static scratch array; log_n at most 20.
extern "C" RustError midnight_hip_fr_device_vec_ntt_zeta_noalloc(void *device_vec, uint32_t log_n, bool inverse,
                                                                 RawFr scale, RawFr first, RawFr second)
{
    if (!device_vec) return make_error((int)hipErrorInvalidValue, "device vector is null");
    DeviceVec *dv = static_cast<DeviceVec *>(device_vec);
    if (log_n >= 63 || dv->len != (size_t(1) << log_n))
        return make_error((int)hipErrorInvalidValue, "device vector length mismatch");
    if (log_n > kNoAllocMaxLog)
        return make_error((int)hipErrorInvalidValue, "transform above 2^20, the no-allocation scratch limit");
    if (log_n == 0) return ok();
    MH_TRY
    Fr *scr = noalloc_scratch();
    hipStream_t st = cur_stream();
    NttJob job;
    job.in = dv->values.get();
    job.out = dv->values.get();
    job.scratch = ntt_scratch_elems(log_n, 1) ? scr : nullptr;
    job.mid_table = false;   // nothing allocated, not even a twiddle table
    job.in_pstride = job.out_pstride = size_t(1) << log_n;
    job.log_n = log_n;
    job.count = 1;
    job.inverse = inverse;
    job.post_zeta = true;
    job.scale = to_hfr(scale);
    job.first = to_hfr(first);
    job.second = to_hfr(second);
    ntt_run(job, st);
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_ntt_zeta_batch(void *device_vec, uint32_t log_n,
                                                                            size_t count, bool inverse, RawFr scale,
                                                                            RawFr first, RawFr second)
{
    if (!device_vec || count == 0) return make_error((int)hipErrorInvalidValue, "invalid resident NTT batch arguments");
    DeviceVec *dv = static_cast<DeviceVec *>(device_vec);
    if (log_n >= 63) return make_error((int)hipErrorInvalidValue, "resident NTT batch size overflow");
    const size_t len = size_t(1) << log_n;
    if (count > SIZE_MAX / len) return make_error((int)hipErrorInvalidValue, "resident NTT batch size overflow");
    if (dv->len != len * count) return make_error((int)hipErrorInvalidValue, "resident NTT batch length mismatch");
    MH_TRY
    if (log_n == 0) {
        // sppark's length-1 NTT is the identity; only the post-processing applies
        hipStream_t st = cur_stream();
        fr_mul_scalar(dv->values.get(), to_fr(scale), count, st);
        MH_HIP(hipStreamSynchronize(st));
        return ok();
    }
    device_ntt_zeta(dv, log_n, count, inverse, scale, first, second, nullptr, 0);
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_device_vec_divide_ntt_zeta(void *device_vec, uint32_t log_n,
                                                                             const RawFr *factors, size_t factor_len,
                                                                             RawFr scale, RawFr first, RawFr second)
{
    if (!device_vec || !factors || factor_len == 0)
        return make_error((int)hipErrorInvalidValue, "invalid resident quotient arguments");
    DeviceVec *dv = static_cast<DeviceVec *>(device_vec);
    if (log_n >= 63 || dv->len != (size_t(1) << log_n))
        return make_error((int)hipErrorInvalidValue, "device vector length mismatch");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_factors(factor_len);
    MH_HIP(hipMemcpyAsync(d_factors.get(), factors, factor_len * 32, hipMemcpyHostToDevice, st));
    device_ntt_zeta(dv, log_n, 1, true, scale, first, second, d_factors.get(), factor_len);
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_hip_fr_device_vec_ntt_multi(void *device_vec, const size_t *offsets,
                                                          const uint32_t *log_ns, size_t count, bool inverse)
{
    if (!device_vec || (count && (!offsets || !log_ns)))
        return make_error((int)hipErrorInvalidValue, "invalid batched NTT arguments");
    DeviceVec *dv = static_cast<DeviceVec *>(device_vec);
    for (size_t i = 0; i < count; ++i) {
        if (log_ns[i] > 22 || offsets[i] > dv->len || (size_t(1) << log_ns[i]) > dv->len - offsets[i])
            return make_error((int)hipErrorInvalidValue, "batched NTT slice exceeds the device vector");
    }
    {
        std::vector<std::pair<size_t, size_t>> spans(count);
        for (size_t i = 0; i < count; ++i) spans[i] = {offsets[i], size_t(1) << log_ns[i]};
        std::sort(spans.begin(), spans.end());
        for (size_t i = 1; i < count; ++i)
            if (spans[i].first < spans[i - 1].first + spans[i - 1].second)
                return make_error((int)hipErrorInvalidValue, "batched NTT slices overlap");
    }
    MH_TRY
    hipStream_t st = cur_stream();
    // one launch per pass for the whole set (runs of consecutive transforms
    // of one size share a descriptor)
    RaggedNtt set;
    set.prepare(dv->values.get(), offsets, log_ns, count, inverse, st);
    set.enqueue(st);
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

// ---------------------------------------------------------------- Fr vector ops

namespace {

template <class F>
RustError host_vector_op(RawFr *values, size_t len, F launch)
{
    if (len == 0) return ok();
    LaneScratch &ls = lane_scratch();
    std::lock_guard<std::mutex> g(ls.mu);
    hipStream_t st = cur_stream();
    ls.a.ensure(len);
    MH_HIP(hipMemcpyAsync(ls.a.get(), values, len * 32, hipMemcpyHostToDevice, st));
    launch(ls.a.get(), st);
    MH_HIP(hipMemcpyAsync(values, ls.a.get(), len * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
}

}  // namespace

// bls12_381_g1_msm.cu:6582-6610: values[i] = 1 / values[i], 0 stays 0
extern "C" RustError midnight_sppark_bls12_381_fr_invert_assign(RawFr *values, size_t len)
{
    if (len == 0) return ok();
    if (!values) return make_error((int)hipErrorInvalidValue, "values is null");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d(len);
    MH_HIP(hipMemcpyAsync(d.get(), values, len * 32, hipMemcpyHostToDevice, st));
    fr_batch_invert(d.get(), len, st);
    MH_HIP(hipMemcpyAsync(values, d.get(), len * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_mul_assign(RawFr *values, RawFr scalar, size_t len)
{
    MH_TRY
    Fr s = to_fr(scalar);
    return host_vector_op(values, len, [&](Fr *d, hipStream_t st) { fr_mul_scalar(d, s, len, st); });
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_mul_assign_zeta(RawFr *values, RawFr first, RawFr second,
                                                                  size_t len)
{
    MH_TRY
    Fr a = to_fr(first), b = to_fr(second);
    return host_vector_op(values, len, [&](Fr *d, hipStream_t st) { fr_mul_zeta(d, a, b, len, st); });
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_mul_assign_repeating(RawFr *values, const RawFr *factors,
                                                                       size_t factor_len, size_t len)
{
    if (len == 0 || factor_len == 0) return ok();
    MH_TRY
    hipStream_t st0 = cur_stream();
    DevBuf<Fr> d_f(factor_len);
    MH_HIP(hipMemcpyAsync(d_f.get(), factors, factor_len * 32, hipMemcpyHostToDevice, st0));
    return host_vector_op(values, len,
                          [&](Fr *d, hipStream_t st) { fr_mul_repeating(d, d_f.get(), factor_len, len, st); });
    MH_CATCH
}
