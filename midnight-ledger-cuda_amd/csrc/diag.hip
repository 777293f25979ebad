// Device measurements used for roofline reporting: the sustained rate of
// v_mad_u64_u32 (the 32x32+64 multiply-add every field product here is
// built from), and the rate of whole Fp / Fr Montgomery products.
#include <algorithm>
#include <chrono>
#include <cstring>

#include "fields.hpp"
#include "fp28.hpp"
#include "fr29.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

// Eight independent multiply-add chains per thread, kept in registers; the
// result is written so the compiler keeps every chain.
__global__ void __launch_bounds__(256) mad_u64_loop(uint64_t *out, uint32_t iters, uint32_t seed)
{
    uint32_t a = seed + threadIdx.x, b = seed ^ (blockIdx.x * 2654435761u);
    uint64_t acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = (uint64_t)(a + k) << 17;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[k] = (uint64_t)(uint32_t)acc[k] * b + acc[k];
        }
    }
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) s ^= acc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <class C>
__global__ void __launch_bounds__(256) field_mul_loop(uint32_t *out, uint32_t iters)
{
    Fe<C> a = Fe<C>::one(), b = Fe<C>::one();
    a.v[0] ^= threadIdx.x;
    b.v[1] ^= blockIdx.x;
    for (uint32_t i = 0; i < iters; ++i) {
        a = mul<C>(a, b);
        b = mul<C>(b, a);
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < C::N; ++k) s ^= a.v[k] ^ b.v[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) fp28_mul_loop(uint32_t *out, uint32_t iters, int square)
{
    Fp28 a = fp28_one(), b = fp28_one();
    a.v[0] ^= threadIdx.x;
    b.v[1] ^= blockIdx.x;
    for (uint32_t i = 0; i < iters; ++i) {
        if (square) {
            a = fp28_sqr(a);
            b = fp28_sqr(b);
        } else {
            a = fp28_mul(a, b);
            b = fp28_mul(b, a);
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 14; ++k) s ^= a.v[k] ^ b.v[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// bucket additions with the point operands in registers (no memory traffic)
__global__ void __launch_bounds__(256) madd_loop(uint32_t *out, uint32_t iters, int radix28)
{
    Affine q;
    q.x = Fp::one();
    q.y = Fp::one();
    q.x.v[0] ^= threadIdx.x * 7919u;
    q.y.v[1] ^= blockIdx.x * 104729u;
    uint32_t s = 0;
    if (radix28) {
        Affine28 q28 = affine28_from_std(q);
        Xyzz28 acc = xyzz28_from_affine(q28);
        acc.zz.v[0] ^= 1;
        for (uint32_t i = 0; i < iters; ++i) {
            acc = xyzz28_add_affine(acc, q28);
            q28.x.v[2] ^= acc.x.v[3] & 1;
        }
#pragma unroll
        for (int k = 0; k < 14; ++k) s ^= acc.x.v[k] ^ acc.zzz.v[k];
    } else {
        Xyzz acc = xyzz_from_affine(q);
        acc.zz.v[0] ^= 1;
        for (uint32_t i = 0; i < iters; ++i) {
            acc = xyzz_add_affine(acc, q);
            q.x.v[2] ^= acc.x.v[3] & 1;
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) s ^= acc.x.v[k] ^ acc.zzz.v[k];
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// two interleaved dependent chains of radix-2^29 Fr products on one wave
__global__ void __launch_bounds__(64) fr29_chain_loop(uint32_t *out, uint32_t iters)
{
    Fr one = Fr::one();
    one.v[0] ^= threadIdx.x;   // still below r
    Fr29 a = fr29_from32(one), b = a;
    b.v[1] ^= 7u;
    for (uint32_t i = 0; i < iters; ++i) {
        a = fr29_mul(a, b);
        b = fr29_mul(b, a);
    }
    out[threadIdx.x] = a.v[0] ^ b.v[0];
}

double time_launch(void (*fn)(hipStream_t), hipStream_t st)
{
    hipEvent_t e0, e1;
    MH_HIP(hipEventCreate(&e0));
    MH_HIP(hipEventCreate(&e1));
    fn(st);   // warm-up
    MH_HIP(hipEventRecord(e0, st));
    fn(st);
    MH_HIP(hipEventRecord(e1, st));
    MH_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms * 1e-3;
}

uint64_t *g_sink = nullptr;
int g_blocks = 0;
constexpr uint32_t kMadIters = 4096;
constexpr uint32_t kMulIters = 64;

void run_mad(hipStream_t st)
{
    hipLaunchKernelGGL(mad_u64_loop, dim3(g_blocks), dim3(256), 0, st, g_sink, kMadIters, 12345u);
}
void run_fp(hipStream_t st)
{
    hipLaunchKernelGGL(field_mul_loop<FpCfg>, dim3(g_blocks), dim3(g_blocks == 1 ? 64 : 256), 0, st,
                       reinterpret_cast<uint32_t *>(g_sink), kMulIters);
}
void run_fr(hipStream_t st)
{
    hipLaunchKernelGGL(field_mul_loop<FrCfg>, dim3(g_blocks), dim3(g_blocks == 1 ? 64 : 256), 0, st,
                       reinterpret_cast<uint32_t *>(g_sink), kMulIters);
}
void run_fr29(hipStream_t st)
{
    hipLaunchKernelGGL(fr29_chain_loop, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t *>(g_sink), kMulIters);
}

}  // namespace

}  // namespace mh

extern "C" RustError midnight_hip_valu_peak(double *mad_u64_per_s, double *clock_hz, int *num_cus)
{
    using namespace mh;
    try {
        hipDeviceProp_t prop;
        MH_HIP(hipGetDeviceProperties(&prop, device_id()));
        g_blocks = prop.multiProcessorCount * 8;
        if (!g_sink) MH_HIP(hipMalloc(reinterpret_cast<void **>(&g_sink), (size_t)g_blocks * 256 * 8));
        hipStream_t st = cur_stream();
        double t = time_launch(run_mad, st);
        double ops = (double)g_blocks * 256 * kMadIters * 16 * 8;
        if (mad_u64_per_s) *mad_u64_per_s = ops / t;
        if (clock_hz) *clock_hz = (double)prop.clockRate * 1e3;
        if (num_cus) *num_cus = prop.multiProcessorCount;
        return ok();
    } catch (const hip_error &e) {
        return make_error(e.code, e.what());
    }
}

namespace mh {
namespace {

// plain streaming copy, 16 B per lane and iteration (grid-stride)
__global__ void __launch_bounds__(256) k_copy16(const uint4 *src, uint4 *dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace
}  // namespace mh

// One device-to-device copy of `bytes` (a multiple of 16) by a plain copy
// kernel (k_copy16), after one untimed pass; returns its rate.  Also the
// reference launch of known size for hardware-counter byte scaling.
extern "C" RustError midnight_hip_copy_bandwidth(size_t bytes, double *bytes_per_s)
{
    using namespace mh;
    if (bytes == 0 || bytes % 16) return make_error((int)hipErrorInvalidValue, "copy size must be a positive multiple of 16");
    try {
        hipStream_t st = cur_stream();
        DevBuf<uint4> a(bytes / 16), b(bytes / 16);
        const size_t n = bytes / 16;
        const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
        MH_HIP(hipMemsetAsync(a.get(), 1, bytes, st));
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, a.get(), b.get(), n);
        MH_HIP(hipGetLastError());
        hipEvent_t e0, e1;
        MH_HIP(hipEventCreate(&e0));
        MH_HIP(hipEventCreate(&e1));
        MH_HIP(hipEventRecord(e0, st));
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, a.get(), b.get(), n);
        MH_HIP(hipGetLastError());
        MH_HIP(hipEventRecord(e1, st));
        MH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        MH_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (bytes_per_s) *bytes_per_s = 2.0 * (double)bytes / (ms * 1e-3);
        return ok();
    } catch (const hip_error &e) {
        return make_error(e.code, e.what());
    }
}

// Whole-product throughput (products per second) for Fp and Fr; with
// blocks = 1 (one wave) it is the latency of a dependent product chain.
extern "C" RustError midnight_hip_field_mul_rate(int blocks, double *fp_per_s, double *fr_per_s)
{
    using namespace mh;
    try {
        hipDeviceProp_t prop;
        MH_HIP(hipGetDeviceProperties(&prop, device_id()));
        if (!g_sink) MH_HIP(hipMalloc(reinterpret_cast<void **>(&g_sink), (size_t)prop.multiProcessorCount * 8 * 256 * 8));
        // g_sink holds one u64 per thread of multiProcessorCount * 8 blocks
        g_blocks = blocks > 0 ? std::min(blocks, prop.multiProcessorCount * 8) : prop.multiProcessorCount * 8;
        if (!g_sink) MH_HIP(hipMalloc(reinterpret_cast<void **>(&g_sink), (size_t)g_blocks * 256 * 8));
        hipStream_t st = cur_stream();
        double per = (double)g_blocks * (g_blocks == 1 ? 64 : 256) * kMulIters * 2;
        if (fp_per_s) *fp_per_s = per / time_launch(run_fp, st);
        if (fr_per_s) *fr_per_s = per / time_launch(run_fr, st);
        return ok();
    } catch (const hip_error &e) {
        return make_error(e.code, e.what());
    }
}

namespace mh {
namespace {
int g_variant = 0;
constexpr uint32_t kMaddIters = 16;
void run_fp28(hipStream_t st)
{
    hipLaunchKernelGGL(fp28_mul_loop, dim3(g_blocks), dim3(256), 0, st, reinterpret_cast<uint32_t *>(g_sink),
                       kMulIters, g_variant);
}
void run_madd(hipStream_t st)
{
    hipLaunchKernelGGL(madd_loop, dim3(g_blocks), dim3(256), 0, st, reinterpret_cast<uint32_t *>(g_sink), kMaddIters,
                       g_variant);
}
}  // namespace
}  // namespace mh

// Throughput of the two Fp representations in isolation, per second over
// the whole device: out[0] Fp28 products, out[1] Fp28 squarings,
// out[2] XYZZ mixed additions (32-bit limbs), out[3] the same in Fp28.
extern "C" RustError midnight_hip_field_rates28(double out[4])
{
    using namespace mh;
    try {
        hipDeviceProp_t prop;
        MH_HIP(hipGetDeviceProperties(&prop, device_id()));
        if (!g_sink)
            MH_HIP(hipMalloc(reinterpret_cast<void **>(&g_sink), (size_t)prop.multiProcessorCount * 8 * 256 * 8));
        g_blocks = prop.multiProcessorCount * 8;
        hipStream_t st = cur_stream();
        const double per_mul = (double)g_blocks * 256 * kMulIters * 2;
        const double per_madd = (double)g_blocks * 256 * kMaddIters;
        g_variant = 0;
        out[0] = per_mul / time_launch(run_fp28, st);
        g_variant = 1;
        out[1] = per_mul / time_launch(run_fp28, st);
        g_variant = 0;
        out[2] = per_madd / time_launch(run_madd, st);
        g_variant = 1;
        out[3] = per_madd / time_launch(run_madd, st);
        return ok();
    } catch (const hip_error &e) {
        return make_error(e.code, e.what());
    }
}

// Latency of one dependent radix-2^29 Fr product on a single wave (the
// product the Poseidon chain issues), in seconds.
extern "C" RustError midnight_hip_fr29_product_latency(double *seconds)
{
    using namespace mh;
    try {
        hipDeviceProp_t prop;
        MH_HIP(hipGetDeviceProperties(&prop, device_id()));
        if (!g_sink)
            MH_HIP(hipMalloc(reinterpret_cast<void **>(&g_sink), (size_t)prop.multiProcessorCount * 8 * 256 * 8));
        const double t = time_launch(run_fr29, cur_stream());
        if (seconds) *seconds = t / (2.0 * kMulIters);
        return ok();
    } catch (const hip_error &e) {
        return make_error(e.code, e.what());
    }
}
