// Grand products: the permutation and lookup arguments' product columns and
// shifted prefix products (bls12_381_g1_msm.cu:1446-1591 fraction / prefix /
// blind-tail kernels, :4303 fr_shifted_prefix_product, :4342
// fr_permutation_products, :4451 / :4545 fr_lookup_products(_device),
// :5738 trace_context_permutation).
//
// The reference divides per row (one field inversion per row) and runs a
// generic prefix product.  Here no row is inverted: with
//   num_i = prod_j (v_j[i] + delta_j beta omega^i + gamma),
//   den_i = prod_j (v_j[i] + beta sigma_j[i] + gamma),
// the product column is
//   z[k] = z[0] * N_{k-1} / D_{k-1} = z[0] * N_{k-1} * S_k / D,
// where N is the prefix product of num, S_k the suffix product of den from
// k, and D = S_0 the whole product of den.  So a set costs one forward and
// one backward multiplicative scan and one inversion of D per set, done on
// the host after a single small read-back for all sets (an inversion is a
// long dependent chain, ~380 products, that one GPU lane would take hundreds
// of microseconds for).  A zero denominator (probability 2^-255 per row)
// makes that row's fraction zero, as the reference's inversion of zero
// does: num is set to zero and den to one.
//
// Scans are tiled: a 256-lane block scans 2048 elements (8 per lane,
// sequentially in registers, then a Hillis-Steele pass over the 256 lane
// totals in LDS), tile totals are scanned by one block per array, and the
// consumer applies the tile prefix while it reads (no separate pass).
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "devvec.hpp"
#include "fields.hpp"
#include "grandproduct.hpp"
#include "rowpow.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr size_t kScanTile = (size_t)kScanBlock * kScanItems;
constexpr int kMaxScanJobs = 8;

struct ScanJob {
    Fr *data;
    Fr *totals;
    size_t len;
    uint32_t rev;   // 1: suffix products (scan from the end)
};

struct ScanJobs {
    ScanJob j[kMaxScanJobs];
};

__device__ __forceinline__ size_t scan_pos(size_t i, size_t len, uint32_t rev) { return rev ? len - 1 - i : i; }

__device__ __forceinline__ void lds_put(uint32_t (*sh)[kScanBlock], int t, const Fr &v)
{
#pragma unroll
    for (int w = 0; w < 8; ++w) sh[w][t] = v.v[w];
}

__device__ __forceinline__ Fr lds_get(uint32_t (*sh)[kScanBlock], int t)
{
    Fr v;
#pragma unroll
    for (int w = 0; w < 8; ++w) v.v[w] = sh[w][t];
    return v;
}

// Inclusive scan of the 256 lane totals of a block; returns the lane's
// exclusive prefix and, through `total`, the block's product.
__device__ __forceinline__ Fr block_scan(uint32_t (*sh)[8][kScanBlock], const Fr &mine, Fr &total)
{
    const int t = threadIdx.x;
    lds_put(sh[0], t, mine);
    __syncthreads();
    int b = 0;
    for (int off = 1; off < kScanBlock; off <<= 1) {
        Fr x = lds_get(sh[b], t);
        if (t >= off) x = lds_get(sh[b], t - off) * x;
        lds_put(sh[b ^ 1], t, x);
        __syncthreads();
        b ^= 1;
    }
    total = lds_get(sh[b], kScanBlock - 1);
    Fr excl = t > 0 ? lds_get(sh[b], t - 1) : Fr::one();
    __syncthreads();   // the LDS may be reused by the caller
    return excl;
}

// Tile-local inclusive scans: block (x = tile, y = job); tile totals to totals[x].
__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(ScanJobs jobs)
{
    __shared__ uint32_t sh[2][8][kScanBlock];
    const ScanJob J = jobs.j[blockIdx.y];
    const size_t tile0 = blockIdx.x * kScanTile;
    if (tile0 >= J.len) return;   // whole block leaves together
    const size_t base = tile0 + (size_t)threadIdx.x * kScanItems;
    Fr v[kScanItems];
    Fr acc = Fr::one();
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = base + j;
        if (i < J.len) {
            const Fr x = load_fr(J.data + scan_pos(i, J.len, J.rev));
            acc = j == 0 ? x : acc * x;
            v[j] = acc;
        }
    }
    Fr total;
    const Fr excl = block_scan(sh, acc, total);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = base + j;
        if (i < J.len) store_fr(J.data + scan_pos(i, J.len, J.rev), threadIdx.x ? excl * v[j] : v[j]);
    }
    if (threadIdx.x == 0) store_fr(J.totals + blockIdx.x, total);
}

// In-place inclusive scan of each job's tile totals (one block per job).
__global__ __launch_bounds__(kScanBlock) void k_scan_totals(ScanJobs jobs)
{
    __shared__ uint32_t sh[2][8][kScanBlock];
    const ScanJob J = jobs.j[blockIdx.x];
    const size_t count = (J.len + kScanTile - 1) / kScanTile;
    Fr carry = Fr::one();
    for (size_t t0 = 0; t0 < count; t0 += kScanTile) {
        const size_t base = t0 + (size_t)threadIdx.x * kScanItems;
        Fr v[kScanItems];
        Fr acc = Fr::one();
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const size_t i = base + j;
            if (i < count) {
                const Fr x = load_fr(J.totals + i);
                acc = j == 0 ? x : acc * x;
                v[j] = acc;
            }
        }
        Fr total;
        const Fr excl = carry * block_scan(sh, acc, total);
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const size_t i = base + j;
            if (i < count) store_fr(J.totals + i, excl * v[j]);
        }
        carry = carry * total;
    }
}

// Product of elements 0..i (forward) or i..len-1 (rev) from a tile-scanned
// array and its scanned totals.
__device__ __forceinline__ Fr scanned(const Fr *data, const Fr *totals, size_t pos, size_t len, uint32_t rev)
{
    const size_t i = rev ? len - 1 - pos : pos;
    const size_t tile = i / kScanTile;
    const Fr local = load_fr(data + pos);
    return tile ? load_fr(totals + tile - 1) * local : local;
}

void run_scans(ScanJobs &jobs, int count, hipStream_t st)
{
    size_t max_len = 0;
    for (int i = 0; i < count; ++i) max_len = std::max(max_len, jobs.j[i].len);
    if (max_len == 0) return;
    const size_t tiles = ceil_div(max_len, kScanTile);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles, count), dim3(kScanBlock), 0, st, jobs);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_totals, dim3(count), dim3(kScanBlock), 0, st, jobs);
    MH_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- permutation

struct PermArgs {
    PermColumnsView cols;
    const Fr *sigmas;
    const Fr *delta_beta;     // delta_j * beta per column
    const Fr *omega_tab;
    uint32_t omega_levels;
    Fr beta, gamma;
    size_t len;
    uint32_t column_count, chunk_len;
    Fr *num, *den;            // [set][len]
    Fr *num_tot, *den_tot;    // [set][tiles]
    size_t tiles;
};

__device__ __forceinline__ const Fr *perm_column(const PermArgs &a, uint32_t j)
{
    const RawColumnRef c = a.cols.refs[j];
    if (c.kind == 0 && c.index < a.cols.fixed_cols) return a.cols.fixed + (size_t)c.index * a.len;
    if (c.kind == 1 && c.index < a.cols.advice_cols) return a.cols.advice + (size_t)c.index * a.len;
    if (c.kind == 2 && c.index < a.cols.instance_cols) return a.cols.instance + (size_t)c.index * a.len;
    return nullptr;
}

// num / den of every row of set blockIdx.y
__global__ __launch_bounds__(256) void k_perm_fraction(PermArgs a)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= a.len) return;
    const uint32_t s = blockIdx.y;
    const uint32_t c0 = s * a.chunk_len;
    const uint32_t c1 = min(c0 + a.chunk_len, a.column_count);
    const Fr wi = omega_pow_tab(i, a.omega_tab, a.omega_levels);
    Fr num = Fr::one(), den = Fr::one();
    for (uint32_t j = c0; j < c1; ++j) {
        const Fr *col = perm_column(a, j);
        if (!col) continue;
        const Fr v = load_fr(col + i) + a.gamma;
        num = num * (v + load_fr(a.delta_beta + j) * wi);
        den = den * (v + a.beta * load_fr(a.sigmas + (size_t)j * a.len + i));
    }
    if (den.is_zero()) {
        den = Fr::one();
        num = Fr::zero();
    }
    store_fr(a.num + (size_t)s * a.len + i, num);
    store_fr(a.den + (size_t)s * a.len + i, den);
}

// lookup l, row i: num = (ci + beta)(ct + gamma), den = (pi + beta)(pt + gamma)
// (a zero denominator gives a zero fraction, as batch inversion leaves 0)
__global__ __launch_bounds__(256) void k_lookup_fraction(PermArgs a, const Fr *ci, const Fr *ct, const Fr *pi,
                                                         const Fr *pt)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= a.len) return;
    const size_t o = (size_t)blockIdx.y * a.len + i;
    Fr num = (load_fr(ci + o) + a.beta) * (load_fr(ct + o) + a.gamma);
    Fr den = (load_fr(pi + o) + a.beta) * (load_fr(pt + o) + a.gamma);
    if (den.is_zero()) {
        den = Fr::one();
        num = Fr::zero();
    }
    store_fr(a.num + o, num);
    store_fr(a.den + o, den);
}

// Per set: N_{u-1}, S_u and D = S_0 (u = last usable row).
__global__ void k_perm_points(PermArgs a, size_t u, uint32_t sets, Fr *points)
{
    const uint32_t s = threadIdx.x;
    if (s >= sets) return;
    const Fr *num = a.num + (size_t)s * a.len, *den = a.den + (size_t)s * a.len;
    const Fr *nt = a.num_tot + (size_t)s * a.tiles, *dt = a.den_tot + (size_t)s * a.tiles;
    store_fr(points + 3 * s + 0, u ? scanned(num, nt, u - 1, a.len, 0) : Fr::one());
    store_fr(points + 3 * s + 1, scanned(den, dt, u, a.len, 1));
    store_fr(points + 3 * s + 2, scanned(den, dt, 0, a.len, 1));
}

// z[k] = init (k = 0), c * N_{k-1} * S_k (k >= 1), blinds on the tail.
__global__ __launch_bounds__(256) void k_perm_finish(PermArgs a, Fr *out, const Fr *consts, const Fr *blinds,
                                                     size_t blinding)
{
    const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (k >= a.len) return;
    const uint32_t s = blockIdx.y;
    Fr z;
    if (k >= a.len - blinding) {
        z = load_fr(blinds + (size_t)s * blinding + (k - (a.len - blinding)));
    } else if (k == 0) {
        z = load_fr(consts + 2 * s);
    } else {
        const Fr *num = a.num + (size_t)s * a.len, *den = a.den + (size_t)s * a.len;
        const Fr *nt = a.num_tot + (size_t)s * a.tiles, *dt = a.den_tot + (size_t)s * a.tiles;
        z = load_fr(consts + 2 * s + 1) * scanned(num, nt, k - 1, a.len, 0) * scanned(den, dt, k, a.len, 1);
    }
    store_fr(out + (size_t)s * a.len + k, z);
}

// out[k] = init (k = 0), init * P_{k-1} (k >= 1) from a forward-scanned array
__global__ __launch_bounds__(256) void k_shift_finish(Fr *out, const Fr *scan, const Fr *totals, size_t len, Fr init)
{
    const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (k >= len) return;
    store_fr(out + k, k == 0 ? init : init * scanned(scan, totals, k - 1, len, 0));
}

struct GpScratch {
    DevBuf<Fr> num, den, num_tot, den_tot, small;
};

GpScratch &gp_scratch()
{
    static GpScratch s[64][kLaneCount];
    return s[device_id()][current_lane()];
}

Fr fr_of(const HostFr &h)
{
    Fr f;
    std::memcpy(f.v, h.l, 32);
    return f;
}

}  // namespace

void permutation_products_device(Fr *out, size_t len, size_t column_count, size_t chunk_len,
                                 const PermColumnsView &cols, const Fr *sigmas, const HostFr *delta_powers,
                                 const HostFr &beta, const HostFr &gamma, const HostFr &omega, const Fr *blinds,
                                 size_t blinding, hipStream_t st)
{
    MH_CHECK(len > 0 && column_count > 0 && chunk_len > 0 && blinding < len, (int)hipErrorInvalidValue,
             "invalid permutation product arguments");
    const size_t sets = ceil_div(column_count, chunk_len);
    MH_CHECK(sets <= 1024 && column_count < (1u << 31), (int)hipErrorInvalidValue,
             "too many permutation columns");
    const size_t tiles = ceil_div(len, kScanTile);
    auto &sc = gp_scratch();
    sc.num.ensure(sets * len);
    sc.den.ensure(sets * len);
    sc.num_tot.ensure(sets * tiles);
    sc.den_tot.ensure(sets * tiles);
    sc.small.ensure(column_count + 3 * sets + 2 * sets);

    // delta_j * beta per column, then (after the scans) the per-set points
    std::vector<HostFr> db(column_count);
    for (size_t j = 0; j < column_count; ++j) db[j] = delta_powers[j] * beta;
    Fr *d_db = sc.small.get();
    Fr *d_points = d_db + column_count;
    Fr *d_consts = d_points + 3 * sets;
    MH_HIP(hipMemcpyAsync(d_db, db.data(), column_count * 32, hipMemcpyHostToDevice, st));
    DevBuf<Fr> tab;
    uint32_t levels = 1;
    build_omega_table(tab, levels, omega, len, st);

    PermArgs a{};
    a.cols = cols;
    a.sigmas = sigmas;
    a.delta_beta = d_db;
    a.omega_tab = tab.get();
    a.omega_levels = levels;
    a.beta = fr_of(beta);
    a.gamma = fr_of(gamma);
    a.len = len;
    a.column_count = (uint32_t)column_count;
    a.chunk_len = (uint32_t)chunk_len;
    a.num = sc.num.get();
    a.den = sc.den.get();
    a.num_tot = sc.num_tot.get();
    a.den_tot = sc.den_tot.get();
    a.tiles = tiles;
    const dim3 rows((unsigned)ceil_div(len, 256), (unsigned)sets);
    hipLaunchKernelGGL(k_perm_fraction, rows, dim3(256), 0, st, a);
    MH_HIP(hipGetLastError());

    for (size_t s0 = 0; s0 < sets; s0 += kMaxScanJobs / 2) {
        ScanJobs jobs{};
        int cnt = 0;
        for (size_t s = s0; s < std::min(sets, s0 + kMaxScanJobs / 2); ++s) {
            jobs.j[cnt++] = ScanJob{a.num + s * len, a.num_tot + s * tiles, len, 0};
            jobs.j[cnt++] = ScanJob{a.den + s * len, a.den_tot + s * tiles, len, 1};
        }
        run_scans(jobs, cnt, st);
    }

    const size_t u = len - blinding - 1;
    hipLaunchKernelGGL(k_perm_points, dim3(1), dim3((unsigned)(64 * ceil_div(sets, 64))), 0, st, a, u,
                       (uint32_t)sets, d_points);
    MH_HIP(hipGetLastError());
    std::vector<HostFr> pts(3 * sets), consts(2 * sets);
    MH_HIP(hipMemcpyAsync(pts.data(), d_points, 3 * sets * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    HostFr init = HostFr::one();
    for (size_t s = 0; s < sets; ++s) {
        const HostFr c = init * pts[3 * s + 2].inverse();   // D = 0 cannot happen: zero rows were replaced
        consts[2 * s] = init;
        consts[2 * s + 1] = c;
        init = u ? c * pts[3 * s] * pts[3 * s + 1] : init;   // z_s[u]
    }
    MH_HIP(hipMemcpyAsync(d_consts, consts.data(), 2 * sets * 32, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_perm_finish, rows, dim3(256), 0, st, a, out, (const Fr *)d_consts, blinds, blinding);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(st));
}

void lookup_products_device(Fr *out, const Fr *ci, const Fr *ct, const Fr *pi, const Fr *pt, size_t len,
                            size_t count, const HostFr &beta, const HostFr &gamma, const Fr *blinds, size_t blinding,
                            hipStream_t st)
{
    MH_CHECK(len > 0 && count > 0 && blinding < len && count <= 65535, (int)hipErrorInvalidValue,
             "invalid lookup product arguments");
    const size_t tiles = ceil_div(len, kScanTile);
    auto &sc = gp_scratch();
    sc.num.ensure(count * len);
    sc.den.ensure(count * len);
    sc.num_tot.ensure(count * tiles);
    sc.den_tot.ensure(count * tiles);
    sc.small.ensure(3 * count + 2 * count);
    Fr *d_points = sc.small.get();
    Fr *d_consts = d_points + 3 * count;
    PermArgs a{};
    a.beta = fr_of(beta);
    a.gamma = fr_of(gamma);
    a.len = len;
    a.num = sc.num.get();
    a.den = sc.den.get();
    a.num_tot = sc.num_tot.get();
    a.den_tot = sc.den_tot.get();
    a.tiles = tiles;
    const dim3 rows((unsigned)ceil_div(len, 256), (unsigned)count);
    hipLaunchKernelGGL(k_lookup_fraction, rows, dim3(256), 0, st, a, ci, ct, pi, pt);
    MH_HIP(hipGetLastError());
    for (size_t s0 = 0; s0 < count; s0 += kMaxScanJobs / 2) {
        ScanJobs jobs{};
        int cnt = 0;
        for (size_t l = s0; l < std::min(count, s0 + kMaxScanJobs / 2); ++l) {
            jobs.j[cnt++] = ScanJob{a.num + l * len, a.num_tot + l * tiles, len, 0};
            jobs.j[cnt++] = ScanJob{a.den + l * len, a.den_tot + l * tiles, len, 1};
        }
        run_scans(jobs, cnt, st);
    }
    // z[k] = N_{k-1} * S_k / D per lookup: read back D (u = 0 gives it twice)
    for (size_t l0 = 0; l0 < count; l0 += 1024) {
        const size_t n = std::min<size_t>(1024, count - l0);
        PermArgs b = a;
        b.num = a.num + l0 * len;
        b.den = a.den + l0 * len;
        b.num_tot = a.num_tot + l0 * tiles;
        b.den_tot = a.den_tot + l0 * tiles;
        hipLaunchKernelGGL(k_perm_points, dim3(1), dim3((unsigned)(64 * ceil_div(n, 64))), 0, st, b, (size_t)0,
                           (uint32_t)n, d_points + 3 * l0);
        MH_HIP(hipGetLastError());
    }
    std::vector<HostFr> pts(3 * count), consts(2 * count);
    MH_HIP(hipMemcpyAsync(pts.data(), d_points, 3 * count * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    for (size_t l = 0; l < count; ++l) {
        consts[2 * l] = HostFr::one();
        consts[2 * l + 1] = pts[3 * l + 2].inverse();   // D != 0: zero rows were replaced
    }
    MH_HIP(hipMemcpyAsync(d_consts, consts.data(), 2 * count * 32, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_perm_finish, rows, dim3(256), 0, st, a, out, (const Fr *)d_consts, blinds, blinding);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(st));
}

void shifted_prefix_product_device(Fr *out, const Fr *in, size_t len, const HostFr &initial, hipStream_t st)
{
    if (len == 0) return;
    const size_t tiles = ceil_div(len, kScanTile);
    auto &sc = gp_scratch();
    sc.num.ensure(len);
    sc.num_tot.ensure(tiles);
    MH_HIP(hipMemcpyAsync(sc.num.get(), in, len * 32, hipMemcpyDeviceToDevice, st));
    ScanJobs jobs{};
    jobs.j[0] = ScanJob{sc.num.get(), sc.num_tot.get(), len, 0};
    run_scans(jobs, 1, st);
    hipLaunchKernelGGL(k_shift_finish, dim3((unsigned)ceil_div(len, 256)), dim3(256), 0, st, out,
                       (const Fr *)sc.num.get(), (const Fr *)sc.num_tot.get(), len, fr_of(initial));
    MH_HIP(hipGetLastError());
}

}  // namespace mh

using namespace mh;

extern "C" RustError midnight_sppark_bls12_381_fr_shifted_prefix_product(RawFr *values, RawFr initial, size_t len)
{
    if (len == 0) return ok();
    if (!values) return make_error((int)hipErrorInvalidValue, "values is null");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d(len);
    MH_HIP(hipMemcpyAsync(d.get(), values, len * 32, hipMemcpyHostToDevice, st));
    shifted_prefix_product_device(d.get(), d.get(), len, to_hfr(initial), st);
    MH_HIP(hipMemcpyAsync(values, d.get(), len * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_fr_permutation_products(
    void **out, const RawFr *columns, const RawFr *permutations, size_t len, size_t column_count, size_t chunk_len,
    RawFr beta, RawFr gamma, RawFr omega, const RawFr *delta_powers, const RawFr *blind_values,
    size_t blinding_factors)
{
    if (out == nullptr || len == 0 || column_count == 0 || chunk_len == 0 || columns == nullptr ||
        permutations == nullptr || delta_powers == nullptr || blinding_factors >= len ||
        (blinding_factors != 0 && blind_values == nullptr))
        return make_error((int)hipErrorInvalidValue, "invalid permutation product arguments");
    *out = nullptr;
    const size_t sets = (column_count + chunk_len - 1) / chunk_len;
    if (column_count > SIZE_MAX / len || sets > SIZE_MAX / len || column_count >= (1u << 31))
        return make_error((int)hipErrorInvalidValue, "permutation product size overflow");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_cols(column_count * len), d_perm(column_count * len), d_blinds;
    DevBuf<RawColumnRef> d_refs(column_count);
    MH_HIP(hipMemcpyAsync(d_cols.get(), columns, column_count * len * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(d_perm.get(), permutations, column_count * len * 32, hipMemcpyHostToDevice, st));
    std::vector<RawColumnRef> refs(column_count);
    for (size_t j = 0; j < column_count; ++j) refs[j] = RawColumnRef{1, (uint32_t)j};
    MH_HIP(hipMemcpyAsync(d_refs.get(), refs.data(), column_count * sizeof(RawColumnRef), hipMemcpyHostToDevice, st));
    if (blinding_factors) {
        d_blinds.alloc(sets * blinding_factors);
        MH_HIP(hipMemcpyAsync(d_blinds.get(), blind_values, sets * blinding_factors * 32, hipMemcpyHostToDevice, st));
    }
    std::vector<HostFr> dp(column_count);
    for (size_t j = 0; j < column_count; ++j) dp[j] = to_hfr(delta_powers[j]);
    PermColumnsView v;
    v.advice = d_cols.get();
    v.advice_cols = (uint32_t)column_count;
    v.refs = d_refs.get();
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = sets * len;
    dv->values.alloc(dv->len);
    permutation_products_device(dv->values.get(), len, column_count, chunk_len, v, d_perm.get(), dp.data(),
                                to_hfr(beta), to_hfr(gamma), to_hfr(omega), d_blinds.get(), blinding_factors, st);
    *out = dv.release();
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:4451-4543
extern "C" RustError midnight_sppark_bls12_381_fr_lookup_products(
    void **out, const RawFr *compressed_inputs, const RawFr *compressed_tables, const RawFr *permuted_inputs,
    const RawFr *permuted_tables, size_t len, size_t lookup_count, RawFr beta, RawFr gamma,
    const RawFr *blind_values, size_t blinding_factors)
{
    if (out == nullptr || len == 0 || lookup_count == 0 || compressed_inputs == nullptr ||
        compressed_tables == nullptr || permuted_inputs == nullptr || permuted_tables == nullptr ||
        blinding_factors >= len || (blinding_factors != 0 && blind_values == nullptr))
        return make_error((int)hipErrorInvalidValue, "invalid lookup product arguments");
    *out = nullptr;
    if (lookup_count > SIZE_MAX / len) return make_error((int)hipErrorInvalidValue, "lookup product size overflow");
    MH_TRY
    hipStream_t st = cur_stream();
    const size_t total = lookup_count * len;
    DevBuf<Fr> d(4 * total), d_blinds(std::max<size_t>(lookup_count * blinding_factors, 1));
    MH_HIP(hipMemcpyAsync(d.get(), compressed_inputs, total * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(d.get() + total, compressed_tables, total * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(d.get() + 2 * total, permuted_inputs, total * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(d.get() + 3 * total, permuted_tables, total * 32, hipMemcpyHostToDevice, st));
    if (blinding_factors)
        MH_HIP(hipMemcpyAsync(d_blinds.get(), blind_values, lookup_count * blinding_factors * 32,
                              hipMemcpyHostToDevice, st));
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = total;
    dv->values.alloc(total);
    lookup_products_device(dv->values.get(), d.get(), d.get() + total, d.get() + 2 * total, d.get() + 3 * total,
                           len, lookup_count, to_hfr(beta), to_hfr(gamma), d_blinds.get(), blinding_factors, st);
    *out = dv.release();
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:4545-4634
extern "C" RustError midnight_sppark_bls12_381_fr_lookup_products_device(void **out, const void *trace_lagrange,
                                                                         size_t len, size_t lookup_count, RawFr beta,
                                                                         RawFr gamma, const RawFr *blind_values,
                                                                         size_t blinding_factors)
{
    if (out == nullptr || len == 0 || lookup_count == 0 || trace_lagrange == nullptr || blinding_factors >= len ||
        (blinding_factors != 0 && blind_values == nullptr))
        return make_error((int)hipErrorInvalidValue, "invalid resident lookup product arguments");
    *out = nullptr;
    if (lookup_count > SIZE_MAX / len)
        return make_error((int)hipErrorInvalidValue, "resident lookup product size overflow");
    const size_t total = lookup_count * len;
    const auto *trace = static_cast<const DeviceVec *>(trace_lagrange);
    if (trace->len != total * 4) return make_error((int)hipErrorInvalidValue, "resident lookup trace length mismatch");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_blinds(std::max<size_t>(lookup_count * blinding_factors, 1));
    if (blinding_factors)
        MH_HIP(hipMemcpyAsync(d_blinds.get(), blind_values, lookup_count * blinding_factors * 32,
                              hipMemcpyHostToDevice, st));
    const Fr *b = trace->values.get();
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = total;
    dv->values.alloc(total);
    lookup_products_device(dv->values.get(), b, b + total, b + 2 * total, b + 3 * total, len, lookup_count,
                           to_hfr(beta), to_hfr(gamma), d_blinds.get(), blinding_factors, st);
    *out = dv.release();
    return ok();
    MH_CATCH
}
