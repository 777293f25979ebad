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
#include <memory>
#include <vector>

#include "devvec.hpp"
#include "fields.hpp"
#include "fr29.hpp"
#include "polyops.hpp"
#include "rowpow.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

constexpr int kBlock = 256;
constexpr size_t kEvalFill = 256 * 4 * 2 * 64;   // lanes that give every SIMD two waves
constexpr int kScanItems = 8;
constexpr size_t kScanTile = (size_t)kBlock * kScanItems;

constexpr uint32_t kPowPoints = 16;
constexpr uint32_t kPowLevels = 3;   // tables for lengths up to 2^33

// the level bases x^(2048^l) of up to kPowPoints points, in the scaled form
// (x 2^261 mod r), from the host
struct PowPoints {
    Fr b[kPowPoints][kPowLevels];
};

// One lane per entry of the power tables of up to kPowPoints points: entry
// j of level l is base_l^j, an 11-bit square-and-multiply chain of
// radix-2^29 products in the scaled form, then one product back to the
// usual form.
__global__ __launch_bounds__(kBlock) void k_omega_tables(Fr *tab, uint32_t levels, PowPoints pts, uint32_t npts)
{
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t per = (size_t)levels * kOmegaTab;
    if (t >= npts * per) return;
    const uint32_t p = (uint32_t)(t / per), r = (uint32_t)(t % per);
    const uint32_t l = r >> kOmegaBits, j = r & (kOmegaTab - 1);
    Fr bsel = pts.b[0][0];
#pragma unroll
    for (uint32_t i = 0; i < kPowPoints; ++i)   // selected, not indexed: the argument stays in registers
#pragma unroll
        for (uint32_t q = 0; q < kPowLevels; ++q)
            if (p == i && l == q) bsel = pts.b[i][q];
    const Fr29 base = fr29_from32(bsel);
    Fr29 acc = s29_one();
    for (int b = kOmegaBits - 1; b >= 0; --b) {
        acc = fr29_mul(acc, acc);
        if ((j >> b) & 1u) acc = fr29_mul(acc, base);
    }
    store_fr(tab + t, s29_to_usual(acc));
}

// partial[(poly * npts + pt) * nblk + block] = sum over the block's rows
__global__ __launch_bounds__(kBlock) void k_eval(const Fr *polys, size_t len, uint32_t chunk, const Fr *tabs,
                                                 uint32_t levels, uint32_t npts, Fr *partial)
{
    __shared__ uint32_t sh[8][kBlock];
    const uint32_t poly = blockIdx.y, pt = blockIdx.z;
    const Fr *p = polys + (size_t)poly * len;
    const Fr *tab = tabs + (size_t)pt * levels * kOmegaTab;
    const size_t start = ((size_t)blockIdx.x * kBlock + threadIdx.x) * chunk;
    Fr acc = Fr::zero();
    if (start < len) {
        const size_t end = std::min(start + (size_t)chunk, len);
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

// partial[query * nblk + block]: query q evaluates polys[qpoly[q]] at
// point qpt[q] (power tables tabs[pt])
__global__ __launch_bounds__(kBlock) void k_eval_queries(const Fr *const *polys, const uint32_t *qpoly,
                                                         const uint32_t *qpt, size_t len, uint32_t chunk,
                                                         const Fr *tabs, uint32_t levels, Fr *partial)
{
    __shared__ uint32_t sh[8][kBlock];
    const uint32_t q = blockIdx.y;
    const Fr *p = polys[qpoly[q]];
    const Fr *tab = tabs + (size_t)qpt[q] * levels * kOmegaTab;
    const size_t start = ((size_t)blockIdx.x * kBlock + threadIdx.x) * chunk;
    Fr acc = Fr::zero();
    if (start < len) {
        const size_t end = std::min(start + (size_t)chunk, len);
        const Fr x = load_fr(tab + 1);
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
        store_fr(partial + (size_t)q * gridDim.x + blockIdx.x, s);
    }
}

// out[q * len + i] = sum over items t of q: coefficients[t] * polys[poly_idx[t]][i]
__global__ __launch_bounds__(kBlock) void k_q_polys(Fr *out, const Fr *const *polys, const uint32_t *offsets,
                                                    const uint32_t *poly_idx, const Fr *coefficients, size_t len,
                                                    size_t q_count)
{
    const size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (id >= len * q_count) return;
    const size_t q = id / len, i = id - q * len;
    Fr acc = Fr::zero();
    for (uint32_t t = offsets[q]; t < offsets[q + 1]; ++t)
        acc = acc + load_fr(polys[poly_idx[t]] + i) * load_fr(coefficients + t);
    store_fr(out + id, acc);
}

// q[i] = p[i + 1], q[len - 1] = 0 (division by X)
__global__ __launch_bounds__(kBlock) void k_shift_down(Fr *q, const Fr *p, size_t len)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < len) store_fr(q + i, i + 1 < len ? load_fr(p + i + 1) : Fr::zero());
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

// the same with up to kArgPolys polynomials and scalars passed by value in
// the kernel arguments (no upload, so no synchronisation on host arrays)
constexpr uint32_t kArgPolys = 16;
struct LincombArgs {
    const Fr *polys[kArgPolys];
    Fr scalars[kArgPolys];
};

__global__ __launch_bounds__(kBlock) void k_lincomb_args(Fr *out, LincombArgs a, uint32_t count, size_t len,
                                                         Fr constant)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= len) return;
    Fr acc = i == 0 ? constant : Fr::zero();
    for (uint32_t j = 0; j < count; ++j) acc = acc + a.scalars[j] * load_fr(a.polys[j] + i);
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

// Coefficients per lane of an evaluation: 16, or fewer (down to 4) while
// the evaluations would leave the GPU short of lanes -- each lane's Horner
// chain is a run of dependent products, so a small evaluation is bound by
// its length, not by issue.
size_t eval_chunk(size_t len, size_t sums)
{
    size_t c = 16;
    while (c > 4 && sums * ceil_div(len, c) < kEvalFill) c /= 2;
    return c;
}

}  // namespace

void omega_tables_device(Fr *tab, uint32_t levels, const HostFr *points, size_t npts, hipStream_t st)
{
    const size_t per = (size_t)levels * kOmegaTab;
    for (size_t p0 = 0; p0 < npts; p0 += kPowPoints) {
        const size_t m = std::min<size_t>(kPowPoints, npts - p0);
        MH_CHECK(levels <= kPowLevels, (int)hipErrorInvalidValue, "polynomial too long for the power tables");
        PowPoints pts{};
        const HostFr k32 = HostFr::from_u64(32);
        for (size_t i = 0; i < m; ++i) {
            HostFr b = points[p0 + i];
            for (uint32_t l = 0; l < levels; ++l) {
                const HostFr sc = b * k32;   // scaled form
                std::memcpy(pts.b[i][l].v, sc.l, 32);
                for (uint32_t q = 0; q < kOmegaBits; ++q) b = b.sqr();
            }
        }
        hipLaunchKernelGGL(k_omega_tables, dim3((unsigned)ceil_div(m * per, kBlock)), dim3(kBlock), 0, st,
                           tab + p0 * per, levels, pts, (uint32_t)m);
        MH_HIP(hipGetLastError());
    }
}

void poly_eval_many(HostFr *out, const Fr *polys, size_t len, size_t count, const HostFr *points, size_t npts,
                    hipStream_t st)
{
    if (len == 0 || count == 0 || npts == 0) return;
    uint32_t levels = 1;
    DevBuf<Fr> all;
    build_omega_tables(all, levels, points, npts, len, st);
    const size_t chunk = eval_chunk(len, count * npts);
    const size_t nblk = ceil_div(len, (size_t)kBlock * chunk);
    DevBuf<Fr> partial(count * npts * nblk);
    hipLaunchKernelGGL(k_eval, dim3((unsigned)nblk, (unsigned)count, (unsigned)npts), dim3(kBlock), 0, st, polys, len,
                       (uint32_t)chunk, (const Fr *)all.get(), levels, (uint32_t)npts, partial.get());
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
    if (polys.size() <= kArgPolys) {
        LincombArgs a{};
        for (size_t j = 0; j < polys.size(); ++j) {
            a.polys[j] = polys[j];
            std::memcpy(a.scalars[j].v, scalars[j].l, 32);
        }
        Fr c;
        std::memcpy(c.v, constant.l, 32);
        hipLaunchKernelGGL(k_lincomb_args, dim3((unsigned)ceil_div(len, kBlock)), dim3(kBlock), 0, st, out, a,
                           (uint32_t)polys.size(), len, c);
        MH_HIP(hipGetLastError());
        return;
    }
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
    if (len == 0) return;
    if (a.is_zero()) {   // division by X: a shift (through a copy, q may alias p)
        DevBuf<Fr> w(len);
        MH_HIP(hipMemcpyAsync(w.get(), p, len * 32, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_shift_down, dim3((unsigned)ceil_div(len, kBlock)), dim3(kBlock), 0, st, q,
                           (const Fr *)w.get(), len);
        MH_HIP(hipGetLastError());
        return;
    }
    if (len < 2) {   // the quotient of a constant is zero
        MH_HIP(hipMemsetAsync(q, 0, 32, st));
        return;
    }
    // the power tables of a and a^-1, back to back (stream-ordered: the
    // temporaries return to this stream's cache, nothing waits here)
    DevBuf<Fr> tabs, w(len);
    uint32_t levels = 1;
    const HostFr pts[2] = {a, a.inverse()};
    build_omega_tables(tabs, levels, pts, 2, len, st);
    const Fr *tab = tabs.get(), *inv_tab = tabs.get() + (size_t)levels * kOmegaTab;
    const unsigned grid = (unsigned)ceil_div(len, kBlock);
    hipLaunchKernelGGL(k_scale_powers, dim3(grid), dim3(kBlock), 0, st, w.get(), p, len, tab, levels);
    MH_HIP(hipGetLastError());
    const size_t tiles = ceil_div(len, kScanTile);
    DevBuf<Fr> totals(tiles);
    hipLaunchKernelGGL(k_suffix_tiles, dim3((unsigned)tiles), dim3(kBlock), 0, st, w.get(), len, totals.get());
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_suffix_totals, dim3(1), dim3(1), 0, st, totals.get(), tiles);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_kate_finish, dim3(grid), dim3(kBlock), 0, st, q, (const Fr *)w.get(),
                       (const Fr *)totals.get(), len, inv_tab, levels);
    MH_HIP(hipGetLastError());
}

void poly_eval_queries(HostFr *out, const std::vector<const Fr *> &polys, size_t len, const uint32_t *qpoly,
                       const uint32_t *qpt, size_t nq, const HostFr *points, size_t npts, hipStream_t st)
{
    if (nq == 0) return;
    if (len == 0) {
        for (size_t q = 0; q < nq; ++q) out[q] = HostFr::zero();
        return;
    }
    uint32_t levels = 1;
    DevBuf<Fr> all;
    build_omega_tables(all, levels, points, npts, len, st);
    DevBuf<const Fr *> d_polys(polys.size());
    DevBuf<uint32_t> d_q(2 * nq);
    MH_HIP(hipMemcpyAsync(d_polys.get(), polys.data(), polys.size() * sizeof(Fr *), hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(d_q.get(), qpoly, nq * 4, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(d_q.get() + nq, qpt, nq * 4, hipMemcpyHostToDevice, st));
    const size_t chunk = eval_chunk(len, nq);
    const size_t nblk = ceil_div(len, (size_t)kBlock * chunk);
    MH_CHECK(nq < 65536, (int)hipErrorInvalidValue, "too many evaluation queries");
    DevBuf<Fr> partial(nq * nblk);
    hipLaunchKernelGGL(k_eval_queries, dim3((unsigned)nblk, (unsigned)nq), dim3(kBlock), 0, st,
                       (const Fr *const *)d_polys.get(), (const uint32_t *)d_q.get(),
                       (const uint32_t *)d_q.get() + nq, len, (uint32_t)chunk, (const Fr *)all.get(), levels,
                       partial.get());
    MH_HIP(hipGetLastError());
    std::vector<HostFr> h(nq * nblk);
    MH_HIP(hipMemcpyAsync(h.data(), partial.get(), h.size() * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    for (size_t q = 0; q < nq; ++q) {
        HostFr sum = HostFr::zero();
        for (size_t b = 0; b < nblk; ++b) sum = sum + h[q * nblk + b];
        out[q] = sum;
    }
}

void poly_q_combine(Fr *out, const std::vector<const Fr *> &polys, size_t len, const uint32_t *offsets,
                    size_t q_count, const uint32_t *poly_idx, const HostFr *coefficients, size_t items,
                    hipStream_t st)
{
    if (len == 0 || q_count == 0) return;
    DevBuf<const Fr *> d_polys(std::max<size_t>(polys.size(), 1));
    DevBuf<uint32_t> d_idx(q_count + 1 + std::max<size_t>(items, 1));
    DevBuf<Fr> d_coef(std::max<size_t>(items, 1));
    if (!polys.empty())
        MH_HIP(hipMemcpyAsync(d_polys.get(), polys.data(), polys.size() * sizeof(Fr *), hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(d_idx.get(), offsets, (q_count + 1) * 4, hipMemcpyHostToDevice, st));
    if (items) {
        MH_HIP(hipMemcpyAsync(d_idx.get() + q_count + 1, poly_idx, items * 4, hipMemcpyHostToDevice, st));
        MH_HIP(hipMemcpyAsync(d_coef.get(), coefficients, items * 32, hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_q_polys, dim3((unsigned)ceil_div(len * q_count, kBlock)), dim3(kBlock), 0, st, out,
                       (const Fr *const *)d_polys.get(), (const uint32_t *)d_idx.get(),
                       (const uint32_t *)d_idx.get() + q_count + 1, (const Fr *)d_coef.get(), len, q_count);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(st));
}

}  // namespace mh

// ---------------------------------------------------------------- C ABI

using namespace mh;

namespace {

std::unique_ptr<DeviceVec> new_devvec(size_t len)
{
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = len;
    dv->values.alloc(std::max<size_t>(len, 1));
    return dv;
}

std::vector<HostFr> host_frs(const RawFr *v, size_t n)
{
    std::vector<HostFr> h(n);
    for (size_t i = 0; i < n; ++i) h[i] = to_hfr(v[i]);
    return h;
}

// the q-polynomial and query arguments shared by both multi-open preparations
const char *check_multiopen(size_t poly_count, const uint32_t *q_offsets, size_t q_offsets_len,
                            const uint32_t *q_poly_indices, size_t q_items_len, size_t point_count,
                            const uint32_t *query_poly_indices, const uint32_t *query_point_indices,
                            size_t query_count, bool resident)
{
    const size_t q_count = q_offsets_len - 1;
    if (q_offsets[0] != 0 || q_offsets[q_count] != q_items_len)
        return resident ? "invalid resident multi-open q offsets" : "invalid multi-open q offsets";
    for (size_t i = 0; i < q_count; ++i)
        if (q_offsets[i] > q_offsets[i + 1])
            return resident ? "non-monotonic resident q offsets" : "non-monotonic multi-open q offsets";
    for (size_t i = 0; i < q_items_len; ++i)
        if (q_poly_indices[i] >= poly_count)
            return resident ? "resident q polynomial index out of range" : "multi-open q polynomial index out of range";
    for (size_t i = 0; i < query_count; ++i)
        if (query_poly_indices[i] >= poly_count || query_point_indices[i] >= point_count)
            return resident ? "resident query index out of range" : "multi-open query index out of range";
    return nullptr;
}

}  // namespace

// bls12_381_g1_msm.cu:3558-3598: out = sum_j scalars[j] * polys[j] (host arrays)
extern "C" RustError midnight_sppark_bls12_381_fr_poly_lincomb(RawFr *out, const RawFr *polys, size_t poly_len,
                                                               size_t poly_count, const RawFr *scalars)
{
    if (poly_len == 0 || poly_count == 0) return ok();
    if (!out || !polys || !scalars) return make_error((int)hipErrorInvalidValue, "invalid polynomial combination");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_polys(poly_len * poly_count), d_out(poly_len);
    MH_HIP(hipMemcpyAsync(d_polys.get(), polys, poly_len * poly_count * 32, hipMemcpyHostToDevice, st));
    std::vector<const Fr *> ptrs(poly_count);
    for (size_t j = 0; j < poly_count; ++j) ptrs[j] = d_polys.get() + j * poly_len;
    poly_lincomb(d_out.get(), ptrs, host_frs(scalars, poly_count), poly_len, HostFr::zero(), st);
    MH_HIP(hipMemcpyAsync(out, d_out.get(), poly_len * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:3234-3278: new device vector sum_j scalars[j] * polys[j]
extern "C" RustError midnight_sppark_bls12_381_fr_device_poly_lincomb(void **out, const void *polys_vec,
                                                                      size_t poly_len, size_t poly_count,
                                                                      const RawFr *scalars, size_t scalar_count)
{
    if (out == nullptr || polys_vec == nullptr || poly_len == 0 || poly_count == 0 || scalar_count != poly_count ||
        scalars == nullptr || poly_count > SIZE_MAX / poly_len)
        return make_error((int)hipErrorInvalidValue, "invalid resident polynomial combination");
    *out = nullptr;
    const auto *polys = static_cast<const DeviceVec *>(polys_vec);
    if (polys->len != poly_len * poly_count)
        return make_error((int)hipErrorInvalidValue, "resident polynomial combination shape mismatch");
    MH_TRY
    hipStream_t st = cur_stream();
    auto dv = new_devvec(poly_len);
    std::vector<const Fr *> ptrs(poly_count);
    for (size_t j = 0; j < poly_count; ++j) ptrs[j] = polys->values.get() + j * poly_len;
    poly_lincomb(dv->values.get(), ptrs, host_frs(scalars, scalar_count), poly_len, HostFr::zero(), st);
    // the handle may be consumed on another lane: it is returned complete
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:3985-4023: out[j] = poly_j(point) for a contiguous set
extern "C" RustError midnight_sppark_bls12_381_fr_device_eval_polys(RawFr *out, const void *device_vec,
                                                                    size_t poly_len, size_t poly_count, RawFr point)
{
    if (out == nullptr || device_vec == nullptr || poly_len == 0 || poly_count == 0 ||
        poly_count > SIZE_MAX / poly_len)
        return make_error((int)hipErrorInvalidValue, "invalid resident polynomial evaluation arguments");
    const auto *polys = static_cast<const DeviceVec *>(device_vec);
    if (polys->len != poly_len * poly_count)
        return make_error((int)hipErrorInvalidValue, "resident polynomial evaluation shape mismatch");
    MH_TRY
    hipStream_t st = cur_stream();
    const HostFr x = to_hfr(point);
    poly_eval_many(reinterpret_cast<HostFr *>(out), polys->values.get(), poly_len, poly_count, &x, 1, st);
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:3600-3721: q polynomials q[k] = sum of coefficient *
// poly over q k's items, and the query evaluations, from host polynomials
extern "C" RustError midnight_sppark_bls12_381_fr_multiopen_prepare(
    RawFr *q_out, RawFr *eval_out, const RawFr *polys, size_t poly_len, size_t poly_count, const uint32_t *q_offsets,
    size_t q_offsets_len, const uint32_t *q_poly_indices, const RawFr *q_coefficients, size_t q_items_len,
    const RawFr *points, size_t point_count, const uint32_t *query_poly_indices, const uint32_t *query_point_indices,
    size_t query_count)
{
    if (poly_len == 0 || poly_count == 0 || q_offsets_len < 2 || q_out == nullptr || eval_out == nullptr ||
        polys == nullptr || q_offsets == nullptr || q_poly_indices == nullptr || q_coefficients == nullptr ||
        points == nullptr || point_count == 0 || query_poly_indices == nullptr || query_point_indices == nullptr)
        return make_error((int)hipErrorInvalidValue, "invalid multi-open preparation arguments");
    if (const char *e = check_multiopen(poly_count, q_offsets, q_offsets_len, q_poly_indices, q_items_len,
                                        point_count, query_poly_indices, query_point_indices, query_count, false))
        return make_error((int)hipErrorInvalidValue, e);
    const size_t q_count = q_offsets_len - 1;
    if (poly_count > SIZE_MAX / poly_len || q_count > SIZE_MAX / poly_len || point_count > SIZE_MAX / poly_len)
        return make_error((int)hipErrorInvalidValue, "multi-open preparation size overflow");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_polys(poly_len * poly_count), d_q(poly_len * q_count);
    MH_HIP(hipMemcpyAsync(d_polys.get(), polys, poly_len * poly_count * 32, hipMemcpyHostToDevice, st));
    std::vector<const Fr *> ptrs(poly_count);
    for (size_t j = 0; j < poly_count; ++j) ptrs[j] = d_polys.get() + j * poly_len;
    poly_q_combine(d_q.get(), ptrs, poly_len, q_offsets, q_count, q_poly_indices,
                   host_frs(q_coefficients, q_items_len).data(), q_items_len, st);
    const auto pts = host_frs(points, point_count);
    poly_eval_queries(reinterpret_cast<HostFr *>(eval_out), ptrs, poly_len, query_poly_indices, query_point_indices,
                      query_count, pts.data(), point_count, st);
    MH_HIP(hipMemcpyAsync(q_out, d_q.get(), poly_len * q_count * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:3723-3871: the same from resident polynomials named by
// (source, index) in three device vectors (0 dynamic, 1 fixed, 2 extra); the
// q polynomials go to the host (q_out) and/or a new device vector
extern "C" RustError midnight_sppark_bls12_381_fr_multiopen_prepare_device(
    RawFr *q_out, void **q_device_out, RawFr *eval_out, const void *dynamic_source, const void *fixed_source,
    const void *extra_source, size_t poly_len, const RawDevicePolyRef *poly_refs, size_t poly_count,
    const uint32_t *q_offsets, size_t q_offsets_len, const uint32_t *q_poly_indices, const RawFr *q_coefficients,
    size_t q_items_len, const RawFr *points, size_t point_count, const uint32_t *query_poly_indices,
    const uint32_t *query_point_indices, size_t query_count)
{
    if (poly_len == 0 || poly_count == 0 || q_offsets_len < 2 || (q_out == nullptr && q_device_out == nullptr) ||
        eval_out == nullptr || dynamic_source == nullptr || fixed_source == nullptr || extra_source == nullptr ||
        poly_refs == nullptr || q_offsets == nullptr || q_poly_indices == nullptr || q_coefficients == nullptr ||
        points == nullptr || point_count == 0 || query_poly_indices == nullptr || query_point_indices == nullptr)
        return make_error((int)hipErrorInvalidValue, "invalid resident multi-open arguments");
    if (q_device_out) *q_device_out = nullptr;
    const DeviceVec *sources[3] = {static_cast<const DeviceVec *>(dynamic_source),
                                   static_cast<const DeviceVec *>(fixed_source),
                                   static_cast<const DeviceVec *>(extra_source)};
    for (size_t i = 0; i < poly_count; ++i) {
        const RawDevicePolyRef r = poly_refs[i];
        if (r.source > 2 || (size_t)r.index >= sources[r.source]->len / poly_len)
            return make_error((int)hipErrorInvalidValue, "resident polynomial reference out of range");
    }
    if (const char *e = check_multiopen(poly_count, q_offsets, q_offsets_len, q_poly_indices, q_items_len,
                                        point_count, query_poly_indices, query_point_indices, query_count, true))
        return make_error((int)hipErrorInvalidValue, e);
    const size_t q_count = q_offsets_len - 1;
    MH_TRY
    hipStream_t st = cur_stream();
    std::vector<const Fr *> ptrs(poly_count);
    for (size_t j = 0; j < poly_count; ++j)
        ptrs[j] = sources[poly_refs[j].source]->values.get() + (size_t)poly_refs[j].index * poly_len;
    auto q = new_devvec(poly_len * q_count);
    poly_q_combine(q->values.get(), ptrs, poly_len, q_offsets, q_count, q_poly_indices,
                   host_frs(q_coefficients, q_items_len).data(), q_items_len, st);
    const auto pts = host_frs(points, point_count);
    poly_eval_queries(reinterpret_cast<HostFr *>(eval_out), ptrs, poly_len, query_poly_indices, query_point_indices,
                      query_count, pts.data(), point_count, st);
    if (q_out) MH_HIP(hipMemcpyAsync(q_out, q->values.get(), poly_len * q_count * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    if (q_device_out) *q_device_out = q.release();
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:3873-3983: f = sum_k scalars[k] * (q_k divided by
// (X - a) for every point a of q_k's set, remainders dropped)
extern "C" RustError midnight_sppark_bls12_381_fr_multiopen_f_poly_device(
    void **out, const void *q_device_vec, size_t poly_len, const uint32_t *point_offsets, size_t point_offsets_len,
    const RawFr *points, size_t point_count, const RawFr *combination_scalars, size_t scalar_count)
{
    if (out == nullptr || q_device_vec == nullptr || poly_len < 2 || point_offsets == nullptr ||
        point_offsets_len < 2 || points == nullptr || combination_scalars == nullptr)
        return make_error((int)hipErrorInvalidValue, "invalid resident f-polynomial arguments");
    *out = nullptr;
    const size_t q_count = point_offsets_len - 1;
    const auto *q_polys = static_cast<const DeviceVec *>(q_device_vec);
    if (q_count != scalar_count || q_count > SIZE_MAX / poly_len || q_polys->len != q_count * poly_len ||
        point_offsets[0] != 0 || point_offsets[q_count] != point_count)
        return make_error((int)hipErrorInvalidValue, "resident f-polynomial shape mismatch");
    for (size_t i = 0; i < q_count; ++i)
        if (point_offsets[i] > point_offsets[i + 1])
            return make_error((int)hipErrorInvalidValue, "non-monotonic f-polynomial point offsets");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> divided(q_count * poly_len);
    MH_HIP(hipMemcpyAsync(divided.get(), q_polys->values.get(), q_count * poly_len * 32, hipMemcpyDeviceToDevice,
                          st));
    std::vector<const Fr *> ptrs(q_count);
    for (size_t k = 0; k < q_count; ++k) {
        Fr *d = divided.get() + k * poly_len;
        for (uint32_t p = point_offsets[k]; p < point_offsets[k + 1]; ++p) poly_kate_divide(d, d, poly_len, to_hfr(points[p]), st);
        ptrs[k] = d;
    }
    auto f = new_devvec(poly_len);
    poly_lincomb(f->values.get(), ptrs, host_frs(combination_scalars, scalar_count), poly_len, HostFr::zero(), st);
    // the handle may be consumed on another lane: it is returned complete
    MH_HIP(hipStreamSynchronize(st));
    *out = f.release();
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:4025-4110: final = sum_k s_k q_k + s_q f; pi = the
// quotient of (final - final(point)) by (X - point): poly_len - 1 values
extern "C" RustError midnight_sppark_bls12_381_fr_multiopen_pi_device(void **out, const void *q_device_vec,
                                                                      const void *f_device_vec, size_t poly_len,
                                                                      const RawFr *combination_scalars,
                                                                      size_t scalar_count, RawFr point)
{
    if (out == nullptr || q_device_vec == nullptr || f_device_vec == nullptr || poly_len < 2 ||
        combination_scalars == nullptr || scalar_count < 2)
        return make_error((int)hipErrorInvalidValue, "invalid resident pi-polynomial arguments");
    *out = nullptr;
    const size_t q_count = scalar_count - 1;
    const auto *q_polys = static_cast<const DeviceVec *>(q_device_vec);
    const auto *f_poly = static_cast<const DeviceVec *>(f_device_vec);
    if (q_count > SIZE_MAX / poly_len || q_polys->len != q_count * poly_len || f_poly->len != poly_len)
        return make_error((int)hipErrorInvalidValue, "resident pi-polynomial shape mismatch");
    MH_TRY
    hipStream_t st = cur_stream();
    std::vector<const Fr *> ptrs(scalar_count);
    for (size_t k = 0; k < q_count; ++k) ptrs[k] = q_polys->values.get() + k * poly_len;
    ptrs[q_count] = f_poly->values.get();
    DevBuf<Fr> fin(poly_len);
    poly_lincomb(fin.get(), ptrs, host_frs(combination_scalars, scalar_count), poly_len, HostFr::zero(), st);
    // the constant term does not enter the quotient, so final(point) need not be subtracted
    poly_kate_divide(fin.get(), fin.get(), poly_len, to_hfr(point), st);
    auto pi = new_devvec(poly_len - 1);
    MH_HIP(hipMemcpyAsync(pi->values.get(), fin.get(), (poly_len - 1) * 32, hipMemcpyDeviceToDevice, st));
    MH_HIP(hipStreamSynchronize(st));
    *out = pi.release();
    return ok();
    MH_CATCH
}
