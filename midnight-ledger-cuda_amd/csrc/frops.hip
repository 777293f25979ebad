// Element-wise Fr kernels behind the reference's vector helpers
// (bls12_381_g1_msm.cu:1188-1223 mul_assign / zeta / repeating,
// :1613-1646 quotient pieces, :1761-1772 scramble-mask add, :1225 and
// :6582 invert_assign).  All are
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

// k_quotient_pieces and k_blind_pieces in one pass, the blinds by value
struct PieceBlinds {
    Fr b[kMaxArgBlinds];
};
__global__ void k_quotient_pieces_blinded(Fr *out, const Fr *q, size_t poly_len, size_t pieces, PieceBlinds bl)
{
    const size_t total = poly_len * pieces;
    for (size_t f = blockIdx.x * (size_t)blockDim.x + threadIdx.x; f < total; f += (size_t)gridDim.x * blockDim.x) {
        const size_t piece = f / poly_len, idx = f % poly_len;
        Fr v;
        if (idx == poly_len - 1) v = piece + 1 < pieces ? bl.b[piece] : Fr::zero();
        else v = load_fr(q + piece * (poly_len - 1) + idx);
        if (idx == 0 && piece > 0) v = v - bl.b[piece - 1];
        store_fr(out + f, v);
    }
}

__global__ void k_fill(Fr *v, Fr x, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        store_fr(v + i, x);
}

// x^(r - 2) = x^-1 for x != 0 (Fermat), left-to-right over the bits of r - 2
__device__ __forceinline__ Fr fr_inverse(const Fr &x)
{
    constexpr uint64_t E[4] = {0xfffffffeffffffffULL, 0x53bda402fffe5bfeULL, 0x3339d80809a1d805ULL,
                               0x73eda753299d7d48ULL};
    Fr r = x;   // bit 254, the top bit of r - 2
    for (int bit = 253; bit >= 0; --bit) {
        r = r * r;
        if ((E[bit >> 6] >> (bit & 63)) & 1) r = r * x;
    }
    return r;
}

// Montgomery's batch inversion over runs of kInvRun elements per lane: one
// Fermat inversion per run and three products per element; zeros stay zero
// (halo2's batch_invert and the reference's 1/0 both give 0).
constexpr uint32_t kInvRun = 32;

__global__ __launch_bounds__(256) void k_batch_invert(Fr *v, Fr *tmp, size_t n)
{
    const size_t lane = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t b = lane * kInvRun;
    if (b >= n) return;
    const size_t e = b + kInvRun < n ? b + kInvRun : n;
    Fr acc = Fr::one();
    for (size_t i = b; i < e; ++i) {
        const Fr x = load_fr(v + i);
        store_fr(tmp + i, acc);
        if (!x.is_zero()) acc = acc * x;
    }
    Fr inv = fr_inverse(acc);
    for (size_t i = e; i-- > b;) {
        const Fr x = load_fr(v + i);
        if (x.is_zero()) continue;
        store_fr(v + i, inv * load_fr(tmp + i));
        inv = inv * x;
    }
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

void fr_quotient_pieces(Fr *out, const Fr *quotient, size_t poly_len, size_t pieces, const Fr *host_blinds,
                        size_t nblinds, hipStream_t st)
{
    MH_CHECK(nblinds + 1 == pieces && nblinds <= kMaxArgBlinds, (int)hipErrorInvalidValue,
             "quotient piece blinds do not match the piece count");
    PieceBlinds bl{};
    for (size_t i = 0; i < nblinds; ++i) bl.b[i] = host_blinds[i];
    hipLaunchKernelGGL(k_quotient_pieces_blinded, grid_for(poly_len * pieces), dim3(kBlock), 0, st, out, quotient,
                       poly_len, pieces, bl);
    MH_HIP(hipGetLastError());
}

void fr_batch_invert(Fr *v, size_t n, hipStream_t st)
{
    if (!n) return;
    DevBuf<Fr> tmp(n);
    hipLaunchKernelGGL(k_batch_invert, dim3((unsigned)ceil_div(ceil_div(n, (size_t)kInvRun), 256)), dim3(256), 0, st,
                       v, tmp.get(), n);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(st));
}

void fr_fill(Fr *v, const Fr &x, size_t n, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_fill, grid_for(n), dim3(kBlock), 0, st, v, x, n);
    MH_HIP(hipGetLastError());
}

}  // namespace mh
