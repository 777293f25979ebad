// Coefficient-form polynomial operations of the KZG opening step
// (bls12_381_g1_msm.cu:3234 fr_device_poly_lincomb, :3985
// fr_device_eval_polys, :3873 / :4025 multi-open f and pi polynomials; CPU
// semantics vendor/midnight-proofs-0.8.1/src/poly/kzg/mod.rs:1978-2570,
// arithmetic eval_polynomial / kate_division / polynomial_inner_product).
//
// Evaluation at a point x: every lane runs Horner over a contiguous chunk of
// coefficients and multiplies the result by x^(chunk start) from a
// two-level power table (about one product per coefficient); block sums go
// to the host, which adds the few hundred partials.
//
// Division by (X - a), dropping the remainder (kate_division): the
// quotient's coefficients are q_i = sum_{j > i} p_j a^(j - i - 1)
// = a^-(i+1) S_{i+1} with S_k = sum_{j >= k} p_j a^j, so the division is one
// product per coefficient, an additive suffix scan, and one product per
// coefficient again (a != 0; the opening points are random challenges).
#include <algorithm>
#include <cstring>
#include <vector>

#include "fields.hpp"
#include "polyops.hpp"
#include "rowpow.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 16;              // coefficients per lane in evaluations
constexpr int kScanItems = 8;
constexpr size_t kScanTile = (size_t)kBlock * kScanItems;

// partial[(poly * npts + pt) * nblk + block] = sum over the block's rows
__global__ __launch_bounds__(kBlock) void k_eval(const Fr *polys, size_t len, const Fr *tabs, uint32_t levels,
                                                 uint32_t npts, Fr *partial)
{
    __shared__ uint32_t sh[8][kBlock];
    const uint32_t poly = blockIdx.y, pt = blockIdx.z;
    const Fr *p = polys + (size_t)poly * len;
    const Fr *tab = tabs + (size_t)pt * levels * kOmegaTab;
    const size_t start = ((size_t)blockIdx.x * kBlock + threadIdx.x) * kChunk;
    Fr acc = Fr::zero();
    if (start < len) {
        const size_t end = std::min(start + (size_t)kChunk, len);
        const Fr x = load_fr(tab + 1);   // x^1
        for (size_t i = end; i-- > start;) acc = acc * x + load_fr(p + i);
        acc = acc * omega_pow_tab(start, tab, levels);
    }
#pragma unroll
    for (int w = 0; w < 8; ++w) sh[w][threadIdx.x] = acc.v[w];
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            Fr a, b;
#pragma unroll
            for (int w = 0; w < 8; ++w) {
                a.v[w] = sh[w][threadIdx.x];
                b.v[w] = sh[w][threadIdx.x + off];
            }
            a = a + b;
#pragma unroll
            for (int w = 0; w < 8; ++w) sh[w][threadIdx.x] = a.v[w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Fr s;
#pragma unroll
        for (int w = 0; w < 8; ++w) s.v[w] = sh[w][0];
        store_fr(partial + ((size_t)poly * npts + pt) * gridDim.x + blockIdx.x, s);
    }
}

// out[i] = constant (i == 0) + sum_j s_j p_j[i]
__global__ __launch_bounds__(kBlock) void k_lincomb(Fr *out, const Fr *const *polys, const Fr *scalars,
                                                    uint32_t count, size_t len, Fr constant)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= len) return;
    Fr acc = i == 0 ? constant : Fr::zero();
    for (uint32_t j = 0; j < count; ++j) acc = acc + load_fr(scalars + j) * load_fr(polys[j] + i);
    store_fr(out + i, acc);
}

// w[j] = p[j] * a^j
__global__ __launch_bounds__(kBlock) void k_scale_powers(Fr *w, const Fr *p, size_t len, const Fr *tab,
                                                         uint32_t levels)
{
    const size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (j >= len) return;
    store_fr(w + j, load_fr(p + j) * omega_pow_tab(j, tab, levels));
}

// Tile-local inclusive additive suffix scan (position i counts from the end)
__global__ __launch_bounds__(kBlock) void k_suffix_tiles(Fr *d, size_t len, Fr *totals)
{
    __shared__ uint32_t sh[2][8][kBlock];
    const size_t tile0 = blockIdx.x * kScanTile;
    const size_t base = tile0 + (size_t)threadIdx.x * kScanItems;
    Fr v[kScanItems];
    Fr acc = Fr::zero();
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = base + j;
        if (i < len) {
            acc = acc + load_fr(d + (len - 1 - i));
            v[j] = acc;
        }
    }
    const int t = threadIdx.x;
#pragma unroll
    for (int w = 0; w < 8; ++w) sh[0][w][t] = acc.v[w];
    __syncthreads();
    int b = 0;
    for (int off = 1; off < kBlock; off <<= 1) {
        Fr x, y;
#pragma unroll
        for (int w = 0; w < 8; ++w) x.v[w] = sh[b][w][t];
        if (t >= off) {
#pragma unroll
            for (int w = 0; w < 8; ++w) y.v[w] = sh[b][w][t - off];
            x = x + y;
        }
#pragma unroll
        for (int w = 0; w < 8; ++w) sh[b ^ 1][w][t] = x.v[w];
        __syncthreads();
        b ^= 1;
    }
    Fr excl = Fr::zero();
    if (t > 0) {
#pragma unroll
        for (int w = 0; w < 8; ++w) excl.v[w] = sh[b][w][t - 1];
    }
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = base + j;
        if (i < len) store_fr(d + (len - 1 - i), v[j] + excl);
    }
    if (t == kBlock - 1) {
        Fr tot;
#pragma unroll
        for (int w = 0; w < 8; ++w) tot.v[w] = sh[b][w][kBlock - 1];
        store_fr(totals + blockIdx.x, tot);
    }
}

// exclusive prefix of the tile totals, one lane (tiles are few)
__global__ void k_suffix_totals(Fr *totals, size_t tiles)
{
    Fr run = Fr::zero();
    for (size_t t = 0; t < tiles; ++t) {
        const Fr v = load_fr(totals + t);
        store_fr(totals + t, run);
        run = run + v;
    }
}

// q[i] = a^-(i+1) * S_{i+1}, i < len - 1; S from the tiled suffix scan
__global__ __launch_bounds__(kBlock) void k_kate_finish(Fr *q, const Fr *s, const Fr *totals, size_t len,
                                                        const Fr *inv_tab, uint32_t levels)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= len) return;
    if (i + 1 == len) {   // the quotient has len - 1 coefficients
        store_fr(q + i, Fr::zero());
        return;
    }
    const size_t k = i + 1;
    const size_t r = len - 1 - k;   // position from the end
    const Fr sk = load_fr(s + k) + load_fr(totals + r / kScanTile);
    store_fr(q + i, sk * omega_pow_tab(k, inv_tab, levels));
}

}  // namespace

void poly_eval_many(HostFr *out, const Fr *polys, size_t len, size_t count, const HostFr *points, size_t npts,
                    hipStream_t st)
{
    if (len == 0 || count == 0 || npts == 0) return;
    std::vector<DevBuf<Fr>> tabs(npts);
    uint32_t levels = 1;
    for (size_t p = 0; p < npts; ++p) build_omega_table(tabs[p], levels, points[p], len, st);
    DevBuf<Fr> all((size_t)npts * levels * kOmegaTab);
    for (size_t p = 0; p < npts; ++p)
        MH_HIP(hipMemcpyAsync(all.get() + p * levels * kOmegaTab, tabs[p].get(), levels * kOmegaTab * 32,
                              hipMemcpyDeviceToDevice, st));
    const size_t nblk = ceil_div(len, (size_t)kBlock * kChunk);
    DevBuf<Fr> partial(count * npts * nblk);
    hipLaunchKernelGGL(k_eval, dim3((unsigned)nblk, (unsigned)count, (unsigned)npts), dim3(kBlock), 0, st, polys, len,
                       (const Fr *)all.get(), levels, (uint32_t)npts, partial.get());
    MH_HIP(hipGetLastError());
    std::vector<HostFr> h(count * npts * nblk);
    MH_HIP(hipMemcpyAsync(h.data(), partial.get(), h.size() * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    for (size_t q = 0; q < count * npts; ++q) {
        HostFr s = HostFr::zero();
        for (size_t b = 0; b < nblk; ++b) s = s + h[q * nblk + b];
        out[q] = s;
    }
}

void poly_lincomb(Fr *out, const std::vector<const Fr *> &polys, const std::vector<HostFr> &scalars, size_t len,
                  const HostFr &constant, hipStream_t st)
{
    MH_CHECK(polys.size() == scalars.size(), (int)hipErrorInvalidValue, "lincomb shape mismatch");
    if (len == 0) return;
    DevBuf<const Fr *> d_polys(std::max<size_t>(polys.size(), 1));
    DevBuf<Fr> d_sc(std::max<size_t>(scalars.size(), 1));
    if (!polys.empty()) {
        MH_HIP(hipMemcpyAsync(d_polys.get(), polys.data(), polys.size() * sizeof(Fr *), hipMemcpyHostToDevice, st));
        MH_HIP(hipMemcpyAsync(d_sc.get(), scalars.data(), scalars.size() * 32, hipMemcpyHostToDevice, st));
    }
    Fr c;
    std::memcpy(c.v, constant.l, 32);
    hipLaunchKernelGGL(k_lincomb, dim3((unsigned)ceil_div(len, kBlock)), dim3(kBlock), 0, st, out, d_polys.get(),
                       (const Fr *)d_sc.get(), (uint32_t)polys.size(), len, c);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(st));   // the host arrays above go out of scope
}

void poly_kate_divide(Fr *q, const Fr *p, size_t len, const HostFr &a, hipStream_t st)
{
    MH_CHECK(!a.is_zero(), (int)hipErrorInvalidValue, "division point is zero");
    if (len < 2) return;
    DevBuf<Fr> tab, inv_tab, w(len);
    uint32_t levels = 1, levels2 = 1;
    build_omega_table(tab, levels, a, len, st);
    build_omega_table(inv_tab, levels2, a.inverse(), len, st);
    const unsigned grid = (unsigned)ceil_div(len, kBlock);
    hipLaunchKernelGGL(k_scale_powers, dim3(grid), dim3(kBlock), 0, st, w.get(), p, len, (const Fr *)tab.get(),
                       levels);
    MH_HIP(hipGetLastError());
    const size_t tiles = ceil_div(len, kScanTile);
    DevBuf<Fr> totals(tiles);
    hipLaunchKernelGGL(k_suffix_tiles, dim3((unsigned)tiles), dim3(kBlock), 0, st, w.get(), len, totals.get());
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_suffix_totals, dim3(1), dim3(1), 0, st, totals.get(), tiles);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_kate_finish, dim3(grid), dim3(kBlock), 0, st, q, (const Fr *)w.get(),
                       (const Fr *)totals.get(), len, (const Fr *)inv_tab.get(), levels2);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(st));
}

}  // namespace mh
