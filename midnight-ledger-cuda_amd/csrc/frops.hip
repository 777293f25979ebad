// Element-wise Fr kernels behind the reference's vector helpers
// (bls12_381_g1_msm.cu:1188-1223 mul_assign / zeta / repeating,
// :1613-1646 quotient pieces, :1761-1772 scramble-mask add).  All are
// HBM-bound streaming passes: one 32-byte read and write per element,
// vectorised as two 16-byte accesses, grid-stride over the vector.
#include "fields.hpp"
#include "frops.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

constexpr int kBlock = 256;

inline dim3 grid_for(size_t n)
{
    size_t g = ceil_div(n, kBlock);
    if (g > 65535u * 32u) g = 65535u * 32u;
    return dim3((unsigned)g);
}

__global__ void k_mul_scalar(Fr *v, Fr s, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        store_fr(v + i, load_fr(v + i) * s);
}

__global__ void k_mul_zeta(Fr *v, Fr first, Fr second, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t c = (uint32_t)(i % 3);
        if (c == 0) continue;
        Fr x = load_fr(v + i);
        store_fr(v + i, x * (c == 1 ? first : second));
    }
}

__global__ void k_mul_repeating(Fr *v, const Fr *f, size_t m, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        store_fr(v + i, load_fr(v + i) * load_fr(f + i % m));
}

__global__ void k_add_mask(Fr *out, const Fr *in, const Fr *mask, size_t m, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        store_fr(out + i, load_fr(in + i) + load_fr(mask + i % m));
}

__global__ void k_quotient_pieces(Fr *out, const Fr *q, size_t poly_len, size_t pieces)
{
    size_t total = poly_len * pieces;
    for (size_t f = blockIdx.x * (size_t)blockDim.x + threadIdx.x; f < total; f += (size_t)gridDim.x * blockDim.x) {
        size_t piece = f / poly_len, idx = f % poly_len;
        Fr v = idx == poly_len - 1 ? Fr::zero() : load_fr(q + piece * (poly_len - 1) + idx);
        store_fr(out + f, v);
    }
}

// piece i gets +blind_i at its last index, piece i+1 gets -blind_i at index 0
__global__ void k_blind_pieces(Fr *pieces, const Fr *blinds, size_t poly_len, size_t count)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i + 1 >= count) return;
    Fr b = load_fr(blinds + i);
    store_fr(pieces + i * poly_len + poly_len - 1, b);
    Fr *next = pieces + (i + 1) * poly_len;
    store_fr(next, load_fr(next) - b);
}

__global__ void k_fill(Fr *v, Fr x, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        store_fr(v + i, x);
}

}  // namespace

void fr_mul_scalar(Fr *v, const Fr &s, size_t n, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_mul_scalar, grid_for(n), dim3(kBlock), 0, st, v, s, n);
    MH_HIP(hipGetLastError());
}

void fr_mul_zeta(Fr *v, const Fr &first, const Fr &second, size_t n, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_mul_zeta, grid_for(n), dim3(kBlock), 0, st, v, first, second, n);
    MH_HIP(hipGetLastError());
}

void fr_mul_repeating(Fr *v, const Fr *factors, size_t m, size_t n, hipStream_t st)
{
    if (!n || !m) return;
    hipLaunchKernelGGL(k_mul_repeating, grid_for(n), dim3(kBlock), 0, st, v, factors, m, n);
    MH_HIP(hipGetLastError());
}

void fr_add_mask(Fr *out, const Fr *in, const Fr *mask, size_t m, size_t n, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_add_mask, grid_for(n), dim3(kBlock), 0, st, out, in, mask, m, n);
    MH_HIP(hipGetLastError());
}

void fr_quotient_pieces(Fr *out, const Fr *quotient, size_t poly_len, size_t pieces, const Fr *blinds,
                        hipStream_t st)
{
    size_t total = poly_len * pieces;
    hipLaunchKernelGGL(k_quotient_pieces, grid_for(total), dim3(kBlock), 0, st, out, quotient, poly_len, pieces);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_blind_pieces, dim3((unsigned)ceil_div(pieces, kBlock)), dim3(kBlock), 0, st, out, blinds,
                       poly_len, pieces);
    MH_HIP(hipGetLastError());
}

void fr_fill(Fr *v, const Fr &x, size_t n, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_fill, grid_for(n), dim3(kBlock), 0, st, v, x, n);
    MH_HIP(hipGetLastError());
}

}  // namespace mh
