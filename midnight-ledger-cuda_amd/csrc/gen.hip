// Device-side construction of point sets: [t_i]G for scalars already in HBM
// (SRS powers [tau^i]G, or random multiples of the generator), by a
// fixed-base comb: T[j][b] = [b * 256^j]G for 32 byte positions j, so that
// [t]G = sum_j T[j][byte_j(t)] costs at most 32 mixed additions, then one
// Fermat inversion per point to return affine coordinates.  Also the
// canonical -> Montgomery conversion of a resident Fr vector.
#include <cstring>
#include <memory>
#include <mutex>

#include "curve.hpp"
#include "devvec.hpp"
#include "frops.hpp"
#include "msm_context.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

// generator, Montgomery form (R = 2^384), 32-bit limbs
__device__ __forceinline__ Affine generator()
{
    constexpr uint32_t X[12] = {0xfd530c16u, 0x5cb38790u, 0x9976fff5u, 0x7817fc67u, 0x143ba1c1u, 0x154f95c7u,
                                0xf3d0e747u, 0xf0ae6acdu, 0x21dbf440u, 0xedce6eccu, 0x9e0bfb75u, 0x12017741u};
    constexpr uint32_t Y[12] = {0x0ce72271u, 0xbaac93d5u, 0x7918fd8eu, 0x8c22631au, 0x570725ceu, 0xdd595f13u,
                                0x50405194u, 0x51ac5829u, 0xad0059c0u, 0x0e1c8c3fu, 0x5008a26au, 0x0bbc3efcu};
    Affine g;
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        g.x.v[i] = X[i];
        g.y.v[i] = Y[i];
    }
    return g;
}

// T[j][b] = [b * 256^j] G, one thread per entry
__global__ void __launch_bounds__(256) k_comb_table(uint32_t *table)
{
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= 32 * 256) return;
    const uint32_t j = id >> 8, b = id & 255;
    Xyzz base = xyzz_from_affine(generator());
    for (uint32_t i = 0; i < 8 * j; ++i) base = xyzz_dbl(base);
    Xyzz acc = Xyzz::inf();
    for (int bit = 7; bit >= 0; --bit) {
        acc = xyzz_dbl(acc);
        if ((b >> bit) & 1u) acc = xyzz_add(acc, base);
    }
    store_affine(table + (uint64_t)id * 24, xyzz_to_affine(acc));
}

__global__ void __launch_bounds__(256) k_fixed_base(const uint32_t *table, const Fr *t, uint32_t *out, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fr s = fr_from_mont(load_fr(t + i));
    Xyzz acc = Xyzz::inf();
    for (uint32_t j = 0; j < 32; ++j) {
        uint32_t b = (s.v[j >> 2] >> (8 * (j & 3))) & 255u;
        if (b) acc = xyzz_add_affine(acc, load_affine(table + ((uint64_t)j * 256 + b) * 24));
    }
    store_affine(out + i * 24, xyzz_to_affine(acc));
}

__global__ void __launch_bounds__(256) k_to_mont(Fr *v, size_t n)
{
    // canonical a -> a * R: Montgomery product with R^2
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr uint32_t R2[8] = {0xf3f29c6du, 0xc999e990u, 0x87925c23u, 0x2b6cedcbu,
                                0x7254398fu, 0x05d31496u, 0x9f59ff11u, 0x0748d9d9u};
    Fr r2;
#pragma unroll
    for (int k = 0; k < 8; ++k) r2.v[k] = R2[k];
    store_fr(v + i, load_fr(v + i) * r2);
}

std::once_flag g_table_once[64];
uint32_t *g_table[64];

const uint32_t *comb_table(hipStream_t st)
{
    int dev = device_id();
    std::call_once(g_table_once[dev], [&]() {
        MH_HIP(hipMalloc(reinterpret_cast<void **>(&g_table[dev]), 32 * 256 * 96));
        hipLaunchKernelGGL(k_comb_table, dim3(32), dim3(256), 0, st, g_table[dev]);
        MH_HIP(hipGetLastError());
        MH_HIP(hipStreamSynchronize(st));
    });
    return g_table[dev];
}

}  // namespace

void generator_multiples(uint32_t *out_points, const Fr *scalars, size_t n, hipStream_t st)
{
    const uint32_t *table = comb_table(st);
    if (!n) return;
    hipLaunchKernelGGL(k_fixed_base, dim3((uint32_t)ceil_div(n, 256)), dim3(256), 0, st, table, scalars, out_points,
                       n);
    MH_HIP(hipGetLastError());
}

}  // namespace mh

using namespace mh;

extern "C" RustError midnight_hip_fr_device_vec_to_montgomery(void *device_vec)
{
    if (!device_vec) return make_error((int)hipErrorInvalidValue, "device vector is null");
    try {
        DeviceVec *dv = static_cast<DeviceVec *>(device_vec);
        hipStream_t st = cur_stream();
        if (dv->len) {
            hipLaunchKernelGGL(k_to_mont, dim3((uint32_t)ceil_div(dv->len, 256)), dim3(256), 0, st,
                               dv->values.get(), dv->len);
            MH_HIP(hipGetLastError());
        }
        MH_HIP(hipStreamSynchronize(st));
        return ok();
    } catch (const hip_error &e) {
        return make_error(e.code, e.what());
    }
}

extern "C" RustError midnight_hip_g1_msm_context_create_from_multiples(void **out, const void *scalar_device_vec)
{
    if (!out || !scalar_device_vec) return make_error((int)hipErrorInvalidValue, "invalid arguments");
    *out = nullptr;
    const DeviceVec *dv = static_cast<const DeviceVec *>(scalar_device_vec);
    if (dv->len >= (size_t(1) << 31)) return make_error((int)hipErrorInvalidValue, "too many MSM points");
    try {
        hipStream_t st = cur_stream();
        std::unique_ptr<MsmContext> ctx = make_context(dv->len);
        generator_multiples(ctx->points.get(), dv->values.get(), dv->len, st);
        build_precomp(*ctx, st);
        MH_HIP(hipStreamSynchronize(st));
        *out = ctx.release();
        return ok();
    } catch (const hip_error &e) {
        return make_error(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return make_error((int)hipErrorOutOfMemory, "host allocation failed");
    }
}

extern "C" RustError midnight_hip_g1_msm_context_points_download(RawG1Affine *out, const void *context, size_t n)
{
    if (!context || (n && !out)) return make_error((int)hipErrorInvalidValue, "invalid arguments");
    const MsmContext *ctx = static_cast<const MsmContext *>(context);
    if (n > ctx->npoints) return make_error((int)hipErrorInvalidValue, "context holds fewer points");
    try {
        hipStream_t st = cur_stream();
        if (n) MH_HIP(hipMemcpyAsync(out, ctx->points.get(), n * 96, hipMemcpyDeviceToHost, st));
        MH_HIP(hipStreamSynchronize(st));
        return ok();
    } catch (const hip_error &e) {
        return make_error(e.code, e.what());
    }
}
