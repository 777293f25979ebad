// BLS12-381 G1 multi-scalar multiplication for gfx950 (signed-window
// Pippenger), batched: one pass of every stage serves all MSMs of a batch.
//
// Stages (all on the lane's stream):
//  1. digits   one thread per (msm, point): Montgomery -> canonical scalar,
//              signed c-bit digits d_w in [-2^(c-1), 2^(c-1)] with carry,
//              histogram of (msm, window, |d|-1) keys by atomics.
//  2. scan     exclusive prefix sums of the histogram -> bucket offsets.
//  3. scatter  digits again; each nonzero digit claims a slot of its bucket
//              (point index | sign << 31).  Order inside a bucket does not
//              matter: the bucket sum is the same group element.
//  4. chunks   every bucket is cut into chunks of at most L entries; chunk
//              offsets by a second scan; a chunk->bucket map.
//  5. accumulate  one lane per chunk: serial XYZZ mixed additions over the
//              chunk's points (8M+2S each), the next point's 96-byte row
//              prefetched while the current one is added.  Single-chunk
//              buckets are written straight to the bucket array.
//  6. bucket sums  one lane per multi-chunk bucket adds its chunk partials.
//  7. bit planes  S_w = sum_i (i+1) B_i = sum_k 2^k Q_k with
//              Q_k = sum of the buckets whose weight has bit k set.  Each Q_k
//              is a plain sum, reduced by two wide GPU passes (groups of G1,
//              then G2), so no GPU lane runs a long serial chain.
//  8. host     the (msm, window, plane) partials come back as Jacobian
//              points; the host sums each plane and runs one doubling chain
//              over the exponents c*w + k (Horner), which is faster on a CPU
//              core than on a single GPU lane.
// The result is the unique group element sum s_i P_i, returned in the
// reference's Jacobian layout (bls12_381_g1_msm.cu:6164-6486 entry points).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "curve.hpp"
#include "devvec.hpp"
#include "fp28.hpp"
#include "frops.hpp"
#include "host_g1.hpp"
#include "msm.hpp"
#include "msm_context.hpp"
#include "profile.hpp"
#include "rngdev.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kSignBit = 0x80000000u;

struct MsmDesc {
    const Fr *scalars;   // device
    uint32_t pt_off;     // first point of this MSM in the point table
    uint32_t n;          // number of (point, scalar) pairs
};

struct Geometry {
    uint32_t c, nwin, B, M;
    // kw: bucket sets per MSM (nwin for plain windows; 1 when the points are
    // precomputed per window, [2^(c w)] P_i at index w * pstride + i)
    uint32_t kw;
    uint32_t pstride;
    uint64_t keys() const { return (uint64_t)M * kw * B; }
};

// Keys are bucket-major: ((b * M) + m) * kw + w.  All threads work on the
// same window at the same time, so a window-major layout would send every
// histogram atomic to one small block of memory (a few L2 channels); with
// the bucket index outermost they spread over the whole counter array.
__device__ __forceinline__ uint64_t key_of(const Geometry &g, uint64_t m_kw, uint32_t w, uint32_t b)
{
    return (uint64_t)b * g.M * g.kw + m_kw + (g.kw > 1 ? w : 0);
}

// signed digit of window w; `carry` flows from window w-1
__device__ __forceinline__ int32_t window_digit(const uint32_t k[9], uint32_t w, uint32_t c, uint32_t &carry)
{
    uint32_t bit = w * c, q = bit >> 5, r = bit & 31;
    uint64_t pair = (uint64_t)k[q] | ((uint64_t)k[q + 1] << 32);
    uint32_t raw = (uint32_t)(pair >> r) & ((1u << c) - 1u);
    raw += carry;
    uint32_t half = 1u << (c - 1);
    if (raw > half) {
        carry = 1;
        return (int32_t)raw - (int32_t)(1u << c);
    }
    carry = 0;
    return (int32_t)raw;
}

// Canonical scalar, replaced by r - s when above (r - 1) / 2 (returns true:
// the point is to be negated), so that it is below 2^254.
__device__ __forceinline__ bool canonical_scalar(const Fr *s, uint32_t k[9])
{
    constexpr uint32_t R[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                               0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
    constexpr uint32_t H[8] = {0x80000000u, 0x7fffffffu, 0x7fff2dffu, 0xa9ded201u,
                               0x04d0ec02u, 0x199cec04u, 0x94cebea4u, 0x39f6d3a9u};   // (r - 1) / 2
    const Fr v = fr_from_mont(load_fr(s));
    // v > H  <=>  H - v borrows
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t d = (uint64_t)H[i] - v.v[i] - borrow;
        borrow = (uint32_t)(d >> 63);
    }
    const bool neg = borrow != 0;
    uint32_t b2 = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t d = (uint64_t)R[i] - v.v[i] - b2;   // r - v
        b2 = (uint32_t)(d >> 63);
        k[i] = neg ? (uint32_t)d : v.v[i];
    }
    k[8] = 0;
    return neg;
}

// Counting sort through LDS.  A block takes a tile of one MSM's scalars and
// one range of that MSM's bucket keys (local key b * kw + w); it counts the
// digits of the tile that fall in the range in an LDS histogram, so global
// memory sees one atomic per (block, bucket) instead of one per digit:
//   pass 0: cnt[key] += count          (bucket sizes)
//   pass 1: base = cursor[key] += count, then each digit takes
//           base + (LDS rank) as its slot.
// Global atomics on this part are device-scope and are resolved past the
// per-XCD L2s, so their count is what sets the sort's speed.
constexpr uint32_t kSortThreads = 1024;

template <uint32_t C>
__global__ void __launch_bounds__(kSortThreads) k_sort_lds(const MsmDesc *descs, Geometry g, uint32_t tile,
                                                           uint32_t range_keys, uint32_t *counters, uint32_t *ent,
                                                           int scatter)
{
    constexpr uint32_t NW = msm_windows(C);
    extern __shared__ uint32_t hist[];
    const uint32_t m = blockIdx.y;
    const MsmDesc d = descs[m];
    const uint32_t i0 = blockIdx.x * tile;
    if (i0 >= d.n) return;   // uniform for the whole block
    const uint32_t i1 = min(i0 + tile, d.n);
    const uint32_t lo = blockIdx.z * range_keys;
    const uint64_t mkw = (uint64_t)m * g.kw;
    for (uint32_t j = threadIdx.x; j < range_keys; j += blockDim.x) hist[j] = 0;
    __syncthreads();
    for (int round = 0; round < (scatter ? 2 : 1); ++round) {
        for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            uint32_t k[9];
            const bool neg = canonical_scalar(d.scalars + i, k);
            uint32_t carry = 0;
#pragma unroll
            for (uint32_t w = 0; w < NW; ++w) {
                int32_t dg = window_digit(k, w, C, carry);
                if (dg == 0) continue;
                uint32_t b = (uint32_t)(dg < 0 ? -dg : dg) - 1;
                uint32_t lkey = b * g.kw + (g.kw > 1 ? w : 0) - lo;
                if (lkey >= range_keys) continue;
                uint32_t r = atomicAdd(&hist[lkey], 1u);
                if (round == 1)
                    ent[r] = (d.pt_off + i + w * g.pstride) | ((dg < 0) != neg ? kSignBit : 0u);
            }
        }
        __syncthreads();
        // flush the histogram: sizes (pass 0) or slot reservation (pass 1)
        for (uint32_t j = threadIdx.x; j < range_keys; j += blockDim.x) {
            uint32_t n = hist[j];
            if (round == 1 || n == 0) continue;
            uint32_t lkey = lo + j;
            uint64_t key = ((uint64_t)(lkey / g.kw) * g.M) * g.kw + mkw + lkey % g.kw;
            uint32_t base = atomicAdd(&counters[key], n);
            if (scatter) hist[j] = base;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- two-pass placement
//
// After the count pass and the scan (off = bucket starts) the digits are
// placed in two passes whose HBM writes are long contiguous runs:
//   pass 1 (k_place_coarse): a block forms the digits of up to 1024
//     scalars, counting-sorts them in LDS by coarse bin (F consecutive keys,
//     at most kCoarseBins bins), reserves its slice of each bin's region of
//     `tmp` -- the same index space as `ent`: bin cb spans off[cb F] ... --
//     with one atomic per bin, and writes its (key, entry) pairs bin by bin;
//   pass 2 (k_place_fine): a block takes up to kFineChunk pairs of one bin,
//     counting-sorts them by key in LDS, reserves each key's slice of its
//     bucket with one atomic, and writes the entries key by key.
// Placing each digit straight into its bucket (one 4-byte store per digit,
// runs of ~8 per block and key) moved ~10x the entry bytes through HBM.
constexpr uint32_t kPlaceThreads = 1024;
constexpr uint32_t kCoarseBins = 1024;   // at most
constexpr uint32_t kStage1 = 16384;      // pairs staged by a pass-1 block
constexpr uint32_t kFineChunk = 16384;   // pairs per pass-2 block
constexpr uint32_t kMaxFine = 4096;      // keys per coarse bin (pass-2 LDS)
constexpr uint32_t kNoKey = 0xffffffffu;

// In-place exclusive scan of a[0, n) in LDS by a block of kPlaceThreads
// (n <= 4 * kPlaceThreads); `wsum` holds 16 wave sums.  Returns the total.
__device__ uint32_t lds_exclusive_scan(uint32_t *a, uint32_t n, uint32_t *wsum)
{
    const uint32_t t = threadIdx.x, per = (n + kPlaceThreads - 1) / kPlaceThreads;
    uint32_t v[4], s = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        const uint32_t idx = t * per + i;
        v[i] = (i < per && idx < n) ? a[idx] : 0u;
        s += v[i];
    }
    uint32_t x = s;   // inclusive scan of s within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if ((int)(t & 63) >= o) x += y;
    }
    const uint32_t wave = t >> 6;
    if ((t & 63) == 63) wsum[wave] = x;
    __syncthreads();
    if (t < 16) {
        uint32_t w = wsum[t];
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const uint32_t y = __shfl_up(w, o, 64);
            if ((int)t >= o) w += y;
        }
        wsum[t] = w;
    }
    __syncthreads();
    uint32_t run = (wave ? wsum[wave - 1] : 0u) + x - s;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        const uint32_t idx = t * per + i;
        if (i < per && idx < n) a[idx] = run;
        run += v[i];
    }
    const uint32_t total = wsum[kPlaceThreads / 64 - 1];
    __syncthreads();
    return total;
}

template <uint32_t C>
__global__ void __launch_bounds__(kPlaceThreads) k_place_coarse(const MsmDesc *descs, Geometry g, uint32_t tile,
                                                               uint32_t fshift, uint32_t nbins, uint32_t *ccursor,
                                                               uint64_t *tmp)
{
    constexpr uint32_t NW = msm_windows(C);
    __shared__ uint32_t hist[kCoarseBins];
    __shared__ uint32_t delta[kCoarseBins];
    __shared__ uint64_t stage[kStage1];
    __shared__ uint32_t wsum[16];
    const uint32_t m = blockIdx.y;
    const MsmDesc d = descs[m];
    const uint32_t i0 = blockIdx.x * tile;
    if (i0 >= d.n) return;   // uniform for the whole block
    const uint32_t t = threadIdx.x;
    for (uint32_t j = t; j < nbins; j += kPlaceThreads) hist[j] = 0;
    __syncthreads();
    const uint32_t i = i0 + t;
    const bool live = t < tile && i < d.n;
    const uint64_t mkw = (uint64_t)m * g.kw;
    uint32_t key[NW], val[NW];
    uint32_t k[9];
    const bool neg = live && canonical_scalar(d.scalars + i, k);
    uint32_t carry = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
        const int32_t dg = live ? window_digit(k, w, C, carry) : 0;
        key[w] = kNoKey;
        if (dg == 0) continue;
        const uint32_t b = (uint32_t)(dg < 0 ? -dg : dg) - 1;
        key[w] = (uint32_t)key_of(g, mkw, w, b);
        val[w] = (d.pt_off + i + w * g.pstride) | ((dg < 0) != neg ? kSignBit : 0u);
        atomicAdd(&hist[key[w] >> fshift], 1u);
    }
    __syncthreads();
    const uint32_t total = lds_exclusive_scan(hist, nbins, wsum);
    for (uint32_t j = t; j < nbins; j += kPlaceThreads) {
        const uint32_t st = hist[j], n = (j + 1 < nbins ? hist[j + 1] : total) - st;
        delta[j] = n ? atomicAdd(&ccursor[j], n) - st : 0u;   // mod 2^32: delta + p is the slot
    }
    __syncthreads();
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
        if (key[w] == kNoKey) continue;
        const uint32_t r = atomicAdd(&hist[key[w] >> fshift], 1u);
        stage[r] = ((uint64_t)key[w] << 32) | val[w];
    }
    __syncthreads();
    for (uint32_t p = t; p < total; p += kPlaceThreads) {
        const uint64_t v = stage[p];
        tmp[delta[(uint32_t)(v >> 32) >> fshift] + p] = v;
    }
}

__global__ void __launch_bounds__(kPlaceThreads) k_place_fine(const uint64_t *tmp, const uint32_t *off,
                                                             const uint32_t *cnt, const uint32_t *cmap2,
                                                             const uint32_t *cst2, const uint32_t *t2,
                                                             uint64_t nkeys, uint32_t fshift, uint32_t *cursor,
                                                             uint32_t *ent)
{
    constexpr uint32_t PER = kFineChunk / kPlaceThreads;
    __shared__ uint32_t fh[kMaxFine];
    __shared__ uint32_t fdelta[kMaxFine];
    __shared__ uint32_t sval[kFineChunk];
    __shared__ uint16_t skey[kFineChunk];
    __shared__ uint32_t wsum[16];
    const uint32_t j = blockIdx.x;
    if (j >= t2[0]) return;   // uniform for the whole block
    const uint32_t t = threadIdx.x;
    const uint32_t cb = cmap2[j];
    const uint32_t r = j - cst2[cb];
    const uint64_t k0 = (uint64_t)cb << fshift;
    const uint32_t nk = (uint32_t)min<uint64_t>(1ull << fshift, nkeys - k0);
    const uint32_t bs = off[k0], be = off[k0 + nk - 1] + cnt[k0 + nk - 1];
    const uint32_t s0 = bs + r * kFineChunk, n = min(be - s0, kFineChunk);
    for (uint32_t x = t; x < nk; x += kPlaceThreads) fh[x] = 0;
    __syncthreads();
    uint64_t v[PER];
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t idx = t + q * kPlaceThreads;
        v[q] = idx < n ? tmp[s0 + idx] : 0ull;
        if (idx < n) atomicAdd(&fh[(uint32_t)(v[q] >> 32) - (uint32_t)k0], 1u);
    }
    __syncthreads();
    const uint32_t total = lds_exclusive_scan(fh, nk, wsum);
    for (uint32_t x = t; x < nk; x += kPlaceThreads) {
        const uint32_t st = fh[x], c = (x + 1 < nk ? fh[x + 1] : total) - st;
        fdelta[x] = c ? atomicAdd(&cursor[k0 + x], c) - st : 0u;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t idx = t + q * kPlaceThreads;
        if (idx >= n) continue;
        const uint32_t fk = (uint32_t)(v[q] >> 32) - (uint32_t)k0;
        const uint32_t rr = atomicAdd(&fh[fk], 1u);
        sval[rr] = (uint32_t)v[q];
        skey[rr] = (uint16_t)fk;
    }
    __syncthreads();
    for (uint32_t p = t; p < n; p += kPlaceThreads) ent[fdelta[skey[p]] + p] = sval[p];
}

#define MH_WINDOW_CASES(X) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)

void launch_digits(bool scatter, uint32_t c, uint64_t max_n, hipStream_t st, const MsmDesc *descs, const Geometry &g,
                   uint32_t *counters, uint32_t *ent)
{
    const uint32_t keys_per_msm = g.kw * g.B;
    // up to 32768 keys (a 128 KiB LDS histogram; gfx950 gives a workgroup
    // 160 KiB): a 16-bit window's 2^15 buckets in one range, so every block
    // derives its digits once
    const uint32_t range_keys = std::min<uint32_t>(keys_per_msm, 32768);
    const uint32_t ranges = (keys_per_msm + range_keys - 1) / range_keys;
    // tiles of up to 16384 scalars (each block flushes one global atomic per
    // touched bucket, so bigger tiles mean fewer atomics), but at least ~512
    // blocks so that a single large MSM still fills the device
    uint64_t tile64 = 16384;
    while (tile64 > 4096 && ceil_div(max_n, tile64) * g.M * ranges < 512) tile64 >>= 1;
    if (max_n < (1u << 18)) tile64 = 4096;
    const uint32_t tile = (uint32_t)tile64;
    dim3 grid((uint32_t)ceil_div(max_n, tile), g.M, ranges);
    const size_t lds = (size_t)range_keys * 4;
    switch (c) {
#define MH_CASE(CC)                                                                                           \
    case CC:                                                                                                  \
        hipLaunchKernelGGL(k_sort_lds<CC>, grid, dim3(kSortThreads), lds, st, descs, g, tile, range_keys,     \
                           counters, ent, scatter ? 1 : 0);                                                   \
        break;
        MH_WINDOW_CASES(MH_CASE)
#undef MH_CASE
    default:
        throw hip_error((int)hipErrorInvalidValue, "unsupported MSM window width");
    }
    MH_HIP(hipGetLastError());
}

void launch_place_coarse(uint32_t c, uint64_t max_n, hipStream_t st, const MsmDesc *descs, const Geometry &g,
                         uint32_t fshift, uint32_t nbins, uint32_t *ccursor, uint64_t *tmp)
{
    const uint32_t nw = msm_windows(c);
    const uint32_t tile = std::min<uint32_t>(kPlaceThreads, kStage1 / nw);
    dim3 grid((uint32_t)ceil_div(max_n, tile), g.M);
    switch (c) {
#define MH_CASE(CC)                                                                                           \
    case CC:                                                                                                  \
        hipLaunchKernelGGL(k_place_coarse<CC>, grid, dim3(kPlaceThreads), 0, st, descs, g, tile, fshift, nbins, \
                           ccursor, tmp);                                                                     \
        break;
        MH_WINDOW_CASES(MH_CASE)
#undef MH_CASE
    default:
        throw hip_error((int)hipErrorInvalidValue, "unsupported MSM window width");
    }
    MH_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- scan

constexpr int kScanItems = 4;
constexpr uint32_t kScanTile = kBlock * kScanItems;

// exclusive scan of one tile per block; block totals to `sums`
__global__ void __launch_bounds__(kBlock) k_scan_tiles(const uint32_t *in, uint32_t *out, uint32_t *sums, uint64_t n)
{
    __shared__ uint32_t s[kBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems], t = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = base + j < n ? in[base + j] : 0u;
        t += v[j];
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t off = 1; off < kBlock; off <<= 1) {
        uint32_t add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - t;   // exclusive prefix of this thread
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == kBlock - 1 && sums) sums[blockIdx.x] = s[kBlock - 1];
}

// Exclusive scan of at most kSmallScan values in one workgroup: each thread
// sums a run of consecutive values, the runs' totals are scanned in LDS, and
// each thread writes its run's prefixes.  One launch where the tiled scan
// takes three (a small proof's bucket and chunk scans).
constexpr uint32_t kSmallScanThreads = 1024;
constexpr uint64_t kSmallScan = 32768;
__global__ void __launch_bounds__(kSmallScanThreads) k_scan_small(const uint32_t *in, uint32_t *out, uint32_t *total,
                                                                 uint64_t n)
{
    __shared__ uint32_t s[kSmallScanThreads];
    const uint32_t per = (uint32_t)((n + kSmallScanThreads - 1) / kSmallScanThreads);
    const uint64_t b0 = (uint64_t)threadIdx.x * per, b1 = min(b0 + per, n);
    uint32_t t = 0;
    for (uint64_t i = b0; i < b1; ++i) t += in[i];
    s[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t off = 1; off < kSmallScanThreads; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - t;
    for (uint64_t i = b0; i < b1; ++i) {
        const uint32_t v = in[i];
        out[i] = run;
        run += v;
    }
    if (threadIdx.x == kSmallScanThreads - 1 && total) *total = s[kSmallScanThreads - 1];
}

// Exclusive prefix of each thread's t over the block (kSmallScanThreads
// threads); *total = the sum.  `s` is the block's LDS array, reusable after.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t t, uint32_t *s, uint32_t *total)
{
    s[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t off = 1; off < kSmallScanThreads; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    const uint32_t r = s[threadIdx.x] - t;
    *total = s[kSmallScanThreads - 1];
    __syncthreads();
    return r;
}

// The planning steps of batches of at most kSmallScan bucket keys, each in
// one workgroup (one launch where the multi-block form takes two to four);
// thread t owns the keys [t per, (t + 1) per), at most kSmallPer, whose
// values it keeps in registers (the loads of a run are issued together).
constexpr uint32_t kSmallPer = kSmallScan / kSmallScanThreads;
__device__ __forceinline__ void small_run(uint64_t n, uint64_t &b0, uint32_t &len)
{
    const uint64_t per = (n + kSmallScanThreads - 1) / kSmallScanThreads;
    b0 = min<uint64_t>((uint64_t)threadIdx.x * per, n);
    len = (uint32_t)(min<uint64_t>(b0 + per, n) - b0);
}

// bucket offsets: off = exclusive scan of cnt, cursor = off
__global__ void __launch_bounds__(kSmallScanThreads) k_offsets_small(const uint32_t *cnt, uint32_t *off,
                                                                    uint32_t *cursor, uint64_t nkeys)
{
    __shared__ uint32_t s[kSmallScanThreads];
    uint64_t b0;
    uint32_t len;
    small_run(nkeys, b0, len);
    uint32_t v[kSmallPer], t = 0;
#pragma unroll
    for (uint32_t i = 0; i < kSmallPer; ++i) {
        v[i] = i < len ? cnt[b0 + i] : 0u;
        t += v[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive(t, s, &total);
#pragma unroll
    for (uint32_t i = 0; i < kSmallPer; ++i) {
        if (i < len) {
            off[b0 + i] = run;
            cursor[b0 + i] = run;
        }
        run += v[i];
    }
}

// chunks: q = ceil(cnt / L), co = exclusive scan of q, the chunk -> bucket
// map, totals[0] entries, totals[1] chunks, totals[3] the fullest bucket's
// chunk count
__global__ void __launch_bounds__(kSmallScanThreads) k_chunk_plan_small(const uint32_t *cnt, const uint32_t *off,
                                                                       uint32_t *q, uint32_t *co, uint32_t *cmap,
                                                                       uint32_t *totals, uint64_t nkeys, uint32_t L)
{
    __shared__ uint32_t s[kSmallScanThreads];
    __shared__ uint32_t smax;
    if (threadIdx.x == 0) smax = 0;
    uint64_t b0;
    uint32_t len;
    small_run(nkeys, b0, len);
    uint32_t v[kSmallPer], t = 0, mx = 0;
#pragma unroll
    for (uint32_t i = 0; i < kSmallPer; ++i) v[i] = i < len ? cnt[b0 + i] : 0u;
#pragma unroll
    for (uint32_t i = 0; i < kSmallPer; ++i) {
        v[i] = (v[i] + L - 1) / L;
        if (i < len) q[b0 + i] = v[i];
        t += v[i];
        mx = max(mx, v[i]);
    }
    __syncthreads();
    atomicMax(&smax, mx);
    uint32_t total;
    uint32_t run = block_exclusive(t, s, &total);
#pragma unroll
    for (uint32_t i = 0; i < kSmallPer; ++i) {
        if (i < len) {
            co[b0 + i] = run;
            for (uint32_t j = 0; j < v[i]; ++j) cmap[run + j] = (uint32_t)(b0 + i);
        }
        run += v[i];
    }
    if (threadIdx.x == 0) {
        totals[0] = off[nkeys - 1] + cnt[nkeys - 1];
        totals[1] = total;
        totals[3] = smax;
    }
}

// one tree level's plan: groups of G partials of the buckets that hold more
// than G at this level (inputs ceil(q / div)), their exclusive offsets lo,
// the lane -> bucket map and totals[2] = lanes of the level
__global__ void __launch_bounds__(kSmallScanThreads) k_level_plan_small(const uint32_t *q, uint64_t nkeys,
                                                                       uint64_t div, uint32_t G, uint32_t *lo,
                                                                       uint32_t *lmap, uint32_t *totals)
{
    __shared__ uint32_t s[kSmallScanThreads];
    // no bucket holds more than G partials at this level (the fullest
    // bucket's chunk count is totals[3]): no groups, nothing else to write
    if ((totals[3] + div - 1) / div <= G) {
        if (threadIdx.x == 0) totals[2] = 0;
        return;
    }
    uint64_t b0;
    uint32_t len;
    small_run(nkeys, b0, len);
    uint32_t v[kSmallPer], t = 0;
#pragma unroll
    for (uint32_t i = 0; i < kSmallPer; ++i) {
        const uint32_t qprev = i < len ? (uint32_t)(((uint64_t)q[b0 + i] + div - 1) / div) : 0u;
        v[i] = qprev > G ? (qprev + G - 1) / G : 0u;
        t += v[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive(t, s, &total);
#pragma unroll
    for (uint32_t i = 0; i < kSmallPer; ++i) {
        if (i < len) {
            lo[b0 + i] = run;
            for (uint32_t j = 0; j < v[i]; ++j) lmap[run + j] = (uint32_t)(b0 + i);
        }
        run += v[i];
    }
    if (threadIdx.x == 0) totals[2] = total;
}

// placement bins (at most kCoarseBins, one per thread): each bin's first
// entry (its coarse cursor), its pass-2 blocks, their exclusive offsets cst2,
// the pass-2 block -> bin map and t2[0] = pass-2 blocks
__global__ void __launch_bounds__(kSmallScanThreads) k_bin_plan(const uint32_t *off, const uint32_t *cnt,
                                                               uint64_t nkeys, uint32_t fshift, uint32_t nbins,
                                                               uint32_t *ccursor, uint32_t *cst2, uint32_t *cmap2,
                                                               uint32_t *t2)
{
    __shared__ uint32_t s[kSmallScanThreads];
    const uint32_t cb = threadIdx.x;
    uint32_t n = 0;
    if (cb < nbins) {
        const uint64_t k0 = (uint64_t)cb << fshift, kl = min<uint64_t>(k0 + (1ull << fshift), nkeys) - 1;
        const uint32_t bs = off[k0], be = off[kl] + cnt[kl];
        ccursor[cb] = bs;
        n = (be - bs + kFineChunk - 1) / kFineChunk;
    }
    uint32_t total;
    const uint32_t o = block_exclusive(n, s, &total);
    if (cb < nbins) {
        cst2[cb] = o;
        for (uint32_t x = 0; x < n; ++x) cmap2[o + x] = cb;
    }
    if (cb == 0) t2[0] = total;
}

__global__ void __launch_bounds__(kBlock) k_scan_add(uint32_t *out, const uint32_t *sums, uint64_t n)
{
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    const uint32_t add = sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < n) out[base + j] += add;
}

// ---------------------------------------------------------------- chunks

// q[b] = chunks of bucket b; totals[3] = max over buckets (one atomic per
// block), read back by eager runs to launch only the tree levels needed
__global__ void __launch_bounds__(kBlock) k_chunk_counts(const uint32_t *cnt, uint32_t *q, uint64_t nkeys, uint32_t L,
                                                         uint32_t *totals)
{
    __shared__ uint32_t smax;
    if (threadIdx.x == 0) smax = 0;
    __syncthreads();
    const uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (b < nkeys) {
        const uint32_t v = (cnt[b] + L - 1) / L;
        q[b] = v;
        atomicMax(&smax, v);
    }
    __syncthreads();
    if (threadIdx.x == 0 && smax) atomicMax(&totals[3], smax);
}

__global__ void k_chunk_map(const uint32_t *q, const uint32_t *co, uint32_t *cmap, uint64_t nkeys)
{
    uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (b >= nkeys) return;
    uint32_t n = q[b], o = co[b];
    for (uint32_t j = 0; j < n; ++j) cmap[o + j] = (uint32_t)b;
}

__global__ void __launch_bounds__(kBlock) k_accumulate(const uint32_t *pts, const uint32_t *ent, const uint32_t *cnt,
                                                       const uint32_t *off, const uint32_t *q, const uint32_t *co,
                                                       const uint32_t *cmap, const uint32_t *totals, uint32_t L,
                                                       uint32_t *part, uint32_t *bucket)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= totals[1]) return;
    const uint32_t b = cmap[j];
    const uint32_t r = j - co[b];
    // the bucket's cnt entries cut into its q = ceil(cnt / L) chunks at
    // equal lengths (+-1, at most L): no short last chunk leaves its lane
    // idle while the others finish
    const uint32_t qb = q[b], cb = cnt[b];
    const uint32_t start = off[b] + (uint32_t)(((uint64_t)r * cb) / qb);
    const uint32_t end = off[b] + (uint32_t)(((uint64_t)(r + 1) * cb) / qb);
    // each point is converted to radix 2^28 (fully reduced: the rare doubling
    // path takes it as canonical) and added in loose coordinates with the
    // digit's sign, as in the precomputed pass
    Xyzz28 acc = xyzz28_inf();
    uint32_t e = ent[start];
    Affine nxt = load_affine(pts + (uint64_t)(e & ~kSignBit) * 24);
    for (uint32_t x = start; x < end; ++x) {
        const Affine cur = nxt;
        const bool neg = e & kSignBit;
        if (x + 1 < end) {
            e = ent[x + 1];
            nxt = load_affine(pts + (uint64_t)(e & ~kSignBit) * 24);
        }
        acc = xyzz28_add_signed_lz(acc, affine28_from_std(cur), neg);
    }
    if (q[b] == 1) store_xyzz28(bucket + (uint64_t)b * kXyzz28Words, acc);
    else store_xyzz28(part + (uint64_t)j * kXyzz28Words, acc);
}

// Same pass over a precomputed table held in radix 2^28 (28 u32 per point,
// rows padded to 32 u32 so that every gathered point is one 128-byte cache
// line; unpadded, six rows in eight straddle two lines): the bucket sum runs
// in Fp28 and is converted once at the end.
constexpr uint32_t kRow28 = 32;

// kLanes = 4: each chunk on a lane quad (small MSMs that cannot fill the GPU)
template <int kLanes>
__global__ void __launch_bounds__(kBlock) k_accumulate28(const uint32_t *pts, const uint32_t *ent, const uint32_t *cnt,
                                                         const uint32_t *off, const uint32_t *q, const uint32_t *co,
                                                         const uint32_t *cmap, const uint32_t *totals, uint32_t L,
                                                         uint32_t *part, uint32_t *bucket)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = t / kLanes, sub = t % kLanes;
    if (j >= totals[1]) return;
    const uint32_t b = cmap[j];
    const uint32_t r = j - co[b];
    // the bucket's cnt entries cut into its q = ceil(cnt / L) chunks at
    // equal lengths (+-1, at most L): no short last chunk leaves its lane
    // idle while the others finish
    const uint32_t qb = q[b], cb = cnt[b];
    const uint32_t start = off[b] + (uint32_t)(((uint64_t)r * cb) / qb);
    const uint32_t end = off[b] + (uint32_t)(((uint64_t)(r + 1) * cb) / qb);
    Xyzz28 acc = xyzz28_inf();
    // the next row is in flight while the current one is added; the
    // digit's sign goes into the addition (a lazy 2p - y)
    uint32_t e = ent[start];
    Affine28 nxt = load_affine28(pts + (uint64_t)(e & ~kSignBit) * kRow28);
    for (uint32_t x = start; x < end; ++x) {
        const Affine28 cur = nxt;
        const bool neg = e & kSignBit;
        if (x + 1 < end) {
            e = ent[x + 1];
            nxt = load_affine28(pts + (uint64_t)(e & ~kSignBit) * kRow28);
        }
        if constexpr (kLanes == 4) acc = xyzz28_add_affine_quad(acc, cur, neg, sub);
        else acc = xyzz28_add_signed_lz(acc, cur, neg);
    }
    if (sub) return;
    // the reduction kernels take the loose radix-2^28 sums as they are
    if (q[b] == 1) store_xyzz28(bucket + (uint64_t)b * kXyzz28Words, acc);
    else store_xyzz28(part + (uint64_t)j * kXyzz28Words, acc);
}

// Bucket reduction.  After the accumulation, bucket b holds q[b] partials
// at part[co[b] + r].  Levels l = 0, 1, ... shrink the buckets that still
// hold more than G partials to ceil(q / G^(l+1)), summing groups of G into
// the other buffer; then one lane per bucket sums the at most G partials
// left.  Each level gives its groups dense lanes (a count per bucket, an
// exclusive scan, a lane -> bucket map): a wave with only a few active
// lanes runs the point addition about three times slower (measured on
// MI355X: 86 us vs 28 us per addition at 4 vs 16+ active lanes), so lanes
// indexed by chunk, most of them idle at the upper levels, would cost more
// than the map.
__device__ __forceinline__ uint32_t level_len(uint32_t q, uint64_t div)
{
    return (uint32_t)(((uint64_t)q + div - 1) / div);
}

// groups of bucket b at the level whose inputs are ceil(q / div) partials
__global__ void k_level_counts(const uint32_t *q, uint64_t nkeys, uint64_t div, uint32_t G, uint32_t *lc)
{
    const uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (b >= nkeys) return;
    const uint32_t qprev = level_len(q[b], div);
    lc[b] = qprev > G ? (qprev + G - 1) / G : 0u;
}

// lane -> bucket map of a level; totals[2] = lanes of the level
__global__ void k_level_map(const uint32_t *lc, const uint32_t *lo, uint64_t nkeys, uint32_t *lmap, uint32_t *totals)
{
    const uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (b >= nkeys) return;
    const uint32_t n = lc[b], o = lo[b];
    for (uint32_t t = 0; t < n; ++t) lmap[o + t] = (uint32_t)b;
    if (b == nkeys - 1) totals[2] = o + n;
}

// One sum per lane (kLanes = 1) or per quad of lanes (kLanes = 4, for
// passes too small to give every SIMD a wave); `sub` is the lane's place
// in its quad and only its first lane stores.
template <int kLanes>
__device__ __forceinline__ Xyzz28 msm_add(const Xyzz28 &p, const Xyzz28 &q, uint32_t sub)
{
    if constexpr (kLanes == 4) return xyzz28_add_quad(p, q, sub);
    else return xyzz28_add_lz(p, q);
}

// the point held by the lane `kMask` places away (xor, kMask < 32), through
// the LDS crossbar's swizzle (no LDS memory)
template <int kMask>
__device__ __forceinline__ Xyzz28 xyzz28_swizzle_xor(const Xyzz28 &p)
{
    constexpr int kPattern = (kMask << 10) | 0x1f;
    Xyzz28 r;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        r.x.v[i] = (uint32_t)__builtin_amdgcn_ds_swizzle((int)p.x.v[i], kPattern);
        r.y.v[i] = (uint32_t)__builtin_amdgcn_ds_swizzle((int)p.y.v[i], kPattern);
        r.zz.v[i] = (uint32_t)__builtin_amdgcn_ds_swizzle((int)p.zz.v[i], kPattern);
        r.zzz.v[i] = (uint32_t)__builtin_amdgcn_ds_swizzle((int)p.zzz.v[i], kPattern);
    }
    return r;
}

// Sum of the n points get(0..n-1) on a group of kLanes lanes: one lane, a
// quad, or (kLanes = 16) four quads, each adding every fourth point, then
// two rounds that add the quads' sums pairwise (lanes 4 and 8 apart): a
// chain of ceil(n / 4) + 2 quad additions instead of n.  `sub` is the lane's
// place in its group; every lane of a group passes the same n, and groups
// leave a kernel only as a whole (the exchanges read within the group).
// Points x with keep(x) false are skipped (empty buckets).
template <int kLanes, class Get, class Keep>
__device__ __forceinline__ Xyzz28 group_sum(uint32_t n, uint32_t sub, Get get, Keep keep)
{
    Xyzz28 acc = xyzz28_inf();
    if constexpr (kLanes == 16) {
        const uint32_t lane = sub & 3;
        for (uint32_t x = sub >> 2; x < n; x += 4)
            if (keep(x)) acc = xyzz28_add_quad(acc, get(x), lane);
        acc = xyzz28_add_quad(acc, xyzz28_swizzle_xor<4>(acc), lane);
        acc = xyzz28_add_quad(acc, xyzz28_swizzle_xor<8>(acc), lane);
    } else {
        for (uint32_t x = 0; x < n; ++x)
            if (keep(x)) acc = msm_add<kLanes>(acc, get(x), sub);
    }
    return acc;
}

template <int kLanes, class Get>
__device__ __forceinline__ Xyzz28 group_sum(uint32_t n, uint32_t sub, Get get)
{
    return group_sum<kLanes>(n, sub, get, [](uint32_t) { return true; });
}

template <int kLanes>
__global__ void __launch_bounds__(kBlock) k_tree_level(const uint32_t *q, const uint32_t *co, const uint32_t *lmap,
                                                       const uint32_t *lo, const uint32_t *totals, uint64_t div,
                                                       uint32_t G, const uint32_t *in, uint32_t *out)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = t / kLanes, sub = t % kLanes;
    if (j >= totals[2]) return;
    const uint32_t b = lmap[j];
    const uint32_t r = j - lo[b];
    const uint32_t qprev = level_len(q[b], div);
    const uint32_t s0 = r * G, s1 = min(s0 + G, qprev);
    const uint64_t base = co[b] + s0;
    const Xyzz28 acc = group_sum<kLanes>(s1 - s0, sub, [&](uint32_t x) {
        return load_xyzz28(in + (base + x) * kXyzz28Words);
    });
    if (sub == 0) store_xyzz28(out + (co[b] + r) * kXyzz28Words, acc);
}

// Bucket b after `levels` levels: the levels that shrank it are the first
// lv, so its partials are in `part` (lv even) or `part2` (lv odd).  Buckets
// of one chunk were written by the accumulation.
template <int kLanes>
__global__ void __launch_bounds__(kBlock) k_tree_final(const uint32_t *q, const uint32_t *co, uint64_t nkeys,
                                                       uint32_t G, uint32_t levels, const uint32_t *part,
                                                       const uint32_t *part2, uint32_t *bucket)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t b = t / kLanes;
    const uint32_t sub = (uint32_t)(t % kLanes);
    if (b >= nkeys) return;
    uint32_t rem = q[b];
    if (rem <= 1) return;
    uint32_t lv = 0;
    while (rem > G && lv < levels) {
        rem = (rem + G - 1) / G;
        ++lv;
    }
    const uint32_t *src = (lv & 1) ? part2 : part;
    const uint64_t base = co[b];
    const Xyzz28 acc = group_sum<kLanes>(rem, sub, [&](uint32_t x) {
        return load_xyzz28(src + (base + x) * kXyzz28Words);
    });
    if (sub == 0) store_xyzz28(bucket + b * kXyzz28Words, acc);
}

// zeroes the bucket counters and the totals
__global__ void k_clear(uint32_t *cnt, uint64_t n, uint32_t *totals)
{
    const uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i0 < 4) totals[i0] = 0;
    for (uint64_t i = i0; i < n; i += (uint64_t)gridDim.x * blockDim.x) cnt[i] = 0;
}

__global__ void k_copy_u32(uint32_t *dst, const uint32_t *src, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// totals[0] = bucket entries, totals[1] = chunks (both scans are exclusive)
__global__ void k_totals(const uint32_t *cnt, const uint32_t *off, const uint32_t *q, const uint32_t *co,
                         uint64_t nkeys, uint32_t *totals)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        totals[0] = off[nkeys - 1] + cnt[nkeys - 1];
        totals[1] = co[nkeys - 1] + q[nkeys - 1];
    }
}

// ---------------------------------------------------------------- bit planes

// Plane k of a window holds the buckets x-1 with bit k of x set, x in [1, B].
// Its e-th member is x = ((e >> k) << (k + 1)) | (1 << k) | (e & (2^k - 1)).
constexpr uint32_t kMaxWindow = 20;   // widest window any path selects (MH_WINDOW_CASES)

struct PlaneLayout {
    uint32_t c;
    uint32_t per_window;                 // chunks per window
    uint32_t first[kMaxWindow + 1];      // first chunk of plane k within a window
    uint32_t members[kMaxWindow];        // members of plane k
    // planes over line sums (two-level weights, below): h low bits, and the
    // number of line sums per MSM; h = 0 for planes over the buckets
    uint32_t h, nlines;
};

// Two-level weights for a precomputed window's one bucket set of B = 2^(c-1)
// buckets: weight x = hi 2^h + lo, and
//   S = sum_x x B_x = sum_lo lo R_lo + 2^h sum_hi hi C_hi,
//   R_lo = sum of the buckets with x mod 2^h = lo   (lo = 1 .. 2^h - 1),
//   C_hi = sum of the buckets with x >> h = hi       (hi = 1 .. H, H = B >> h),
// so each bucket is added twice (into its R and its C line) instead of
// into ~c/2 bit planes; the planes then run over the 2^h - 1 + H line sums
// (plane k < h over the R lines with bit k of lo, plane h + k over the C
// lines with bit k of hi), and the host's sum_k 2^k Q_k is unchanged.
struct LineLayout {
    uint32_t h, H, nlines;
    uint32_t chunks_r, chunks_c;         // chunks of G1 members per R line, per full C line
    uint32_t per_msm;                    // chunks per MSM
};

template <int kLanes>
__global__ void __launch_bounds__(kBlock) k_plane_partials(const uint32_t *bucket, const uint32_t *cnt, Geometry g,
                                                           PlaneLayout pl, uint32_t G1, uint64_t total, uint32_t *pp)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t id = t / kLanes;
    const uint32_t sub = (uint32_t)(t % kLanes);
    if (id >= total) return;
    uint64_t mw = id / pl.per_window;
    uint32_t rem = (uint32_t)(id % pl.per_window);
    uint32_t k = 0;
    while (rem >= pl.first[k + 1]) ++k;
    uint32_t chunk = rem - pl.first[k];
    uint32_t e0 = chunk * G1, e1 = min(e0 + G1, pl.members[k]);
    const uint64_t stride = (uint64_t)g.M * g.kw;   // keys are bucket-major
    Xyzz28 acc;
    if (pl.h) {
        // over the line sums of MSM mw: R lines 0 .. 2^h - 2, then C lines
        const uint32_t kk = k < pl.h ? k : k - pl.h;
        const uint32_t lowmask = (1u << kk) - 1u;
        const uint64_t base = mw * pl.nlines + (k < pl.h ? 0u : (1u << pl.h) - 1u);
        acc = group_sum<kLanes>(e1 - e0, sub, [&](uint32_t i) {
            const uint32_t e = e0 + i;
            const uint32_t x = ((e >> kk) << (kk + 1)) | (1u << kk) | (e & lowmask);
            return load_xyzz28(bucket + (base + x - 1) * kXyzz28Words);
        });
    } else {
        const uint32_t lowmask = (1u << k) - 1u;
        const auto key = [&](uint32_t i) {
            const uint32_t e = e0 + i;
            const uint32_t x = ((e >> k) << (k + 1)) | (1u << k) | (e & lowmask);
            return (uint64_t)(x - 1) * stride + mw;
        };
        // empty buckets are not written
        acc = group_sum<kLanes>(
            e1 - e0, sub, [&](uint32_t i) { return load_xyzz28(bucket + key(i) * kXyzz28Words); },
            [&](uint32_t i) { return cnt[key(i)] != 0; });
    }
    if (sub == 0) store_xyzz28(pp + id * kXyzz28Words, acc);
}

// Partial sums of the line sums (LineLayout) of a precomputed window's
// buckets: chunk `id` adds at most G1 buckets of one R or C line.
template <int kLanes>
__global__ void __launch_bounds__(kBlock) k_line_partials(const uint32_t *bucket, const uint32_t *cnt, Geometry g,
                                                          LineLayout ll, uint32_t G1, uint64_t total, uint32_t *pp)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t id = t / kLanes;
    const uint32_t sub = (uint32_t)(t % kLanes);
    if (id >= total) return;
    const uint64_t m = id / ll.per_msm;
    const uint32_t rem = (uint32_t)(id % ll.per_msm);
    const uint32_t nr = ((1u << ll.h) - 1u) * ll.chunks_r;
    uint32_t x0, step, e0, e1;   // weights x0 + step * e, e in [e0, e1)
    if (rem < nr) {
        const uint32_t lo = rem / ll.chunks_r + 1, chunk = rem % ll.chunks_r;
        x0 = lo;
        step = 1u << ll.h;
        e0 = chunk * G1;
        e1 = min(e0 + G1, ll.H);
    } else if (rem - nr < (ll.H - 1) * ll.chunks_c) {
        const uint32_t r2 = rem - nr;
        const uint32_t hi = r2 / ll.chunks_c + 1, chunk = r2 % ll.chunks_c;
        x0 = hi << ll.h;
        step = 1;
        e0 = chunk * G1;
        e1 = min(e0 + G1, 1u << ll.h);
    } else {   // C_H: the one bucket of weight B
        x0 = ll.H << ll.h;
        step = 1;
        e0 = 0;
        e1 = 1;
    }
    const uint64_t stride = (uint64_t)g.M * g.kw;   // keys are bucket-major
    const auto key = [&](uint32_t i) { return (uint64_t)(x0 + step * (e0 + i) - 1) * stride + m; };
    // empty buckets are not written
    const Xyzz28 acc = group_sum<kLanes>(
        e1 - e0, sub, [&](uint32_t i) { return load_xyzz28(bucket + key(i) * kXyzz28Words); },
        [&](uint32_t i) { return cnt[key(i)] != 0; });
    if (sub == 0) store_xyzz28(pp + id * kXyzz28Words, acc);
}

// Sum groups of G consecutive partials within segments of `seg` partials;
// output segment length ceil(seg / G).  Results as Jacobian when `to_jac`;
// with `totals`, the batch totals (4 words) are copied after the results, so
// that one read-back brings both.
template <int kLanes>
__global__ void __launch_bounds__(kBlock) k_group_sums(const uint32_t *in, uint64_t nseg, const uint32_t *seg_off,
                                                       const uint32_t *seg_len, const uint32_t *out_off, uint32_t G,
                                                       uint64_t total_out, uint32_t *out, int to_jac,
                                                       const uint32_t *totals)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t id = t / kLanes;
    const uint32_t sub = (uint32_t)(t % kLanes);
    if (id >= total_out) return;
    // locate segment by binary search over out_off (nseg small)
    uint64_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) / 2;
        if (out_off[mid] <= id) lo = mid;
        else hi = mid;
    }
    uint32_t local = (uint32_t)(id - out_off[lo]);
    uint32_t s0 = seg_off[lo] + local * G, s1 = min(s0 + G, seg_off[lo] + seg_len[lo]);
    const Xyzz28 acc = group_sum<kLanes>(s1 - s0, sub, [&](uint32_t x) {
        return load_xyzz28(in + (uint64_t)(s0 + x) * kXyzz28Words);
    });
    if (sub) return;
    // the last pass converts (fully reduced) to the 32-bit Jacobian form
    if (totals && id == 0)
        for (int i = 0; i < 4; ++i) out[total_out * 36 + i] = totals[i];
    if (to_jac) store_jacobian(out + id * 36, xyzz_to_jacobian(xyzz28_to_std(acc)));
    else store_xyzz28(out + id * kXyzz28Words, acc);
}

// row w * n + i of `out` = [2^(c w)] P_i, by c doublings per window, stored
// in radix 2^28 (Affine28) for the accumulation pass
__global__ void __launch_bounds__(kBlock) k_precompute(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t c,
                                                       uint32_t nwin)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Affine a = load_affine(in + (uint64_t)i * 24);
    store_affine28(out + (uint64_t)i * kRow28, affine28_from_std(a));
    Xyzz x = xyzz_from_affine(a);
    for (uint32_t w = 1; w < nwin; ++w) {
        for (uint32_t k = 0; k < c; ++k) x = xyzz_dbl(x);
        store_affine28(out + ((uint64_t)w * n + i) * kRow28, affine28_from_std(xyzz_to_affine(x)));
    }
}

inline uint32_t grid1(uint64_t n) { return (uint32_t)ceil_div(n, kBlock); }

// lanes that give every SIMD of the 256 CUs two waves
constexpr uint64_t kFillLanes = 256 * 4 * 2 * 64;

// items per lane for `work` serial items: `cap`, or fewer (at least 4) when
// `cap` per lane would leave the GPU short of kFillLanes lanes
inline uint32_t fill_group(uint64_t work, uint32_t cap)
{
    const uint64_t g = ceil_div(work, kFillLanes);
    return (uint32_t)std::max<uint64_t>(std::min<uint64_t>(g, cap), std::min<uint32_t>(4, cap));
}

}  // namespace

namespace {

// the table window of the fixed rule: c = log2 n - 4, at most 16: wider
// windows cut the entries (n * ceil(256/c)) but multiply the digit passes of
// the LDS sort (2^(c-1) / 32768 key ranges) and the bucket reduction; at
// 2^24 points c = 20 spent as long sorting and reducing as accumulating.
uint32_t rule_table_window(size_t n)
{
    int lg = 0;
    while ((size_t(1) << lg) < n) ++lg;
    return (uint32_t)std::min(16, std::max(10, lg - 4));
}

// the per-window table of window c into `table`; false when its entries
// would not fit 31-bit indices
bool build_table(const MsmContext &ctx, uint32_t c, DevBuf<uint32_t> &table, hipStream_t st, size_t rows = 0)
{
    const size_t n = rows ? rows : ctx.npoints;
    const uint32_t nwin = msm_windows(c);
    if ((uint64_t)nwin * n >= (1ull << 31)) return false;   // entry indices are 31-bit
    table.alloc((size_t)nwin * n * kRow28);   // Affine28 rows
    hipLaunchKernelGGL(k_precompute, dim3(grid1(n)), dim3(kBlock), 0, st, ctx.points.get(), table.get(),
                       (uint32_t)n, c, nwin);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(st));
    return true;
}

// milliseconds of one batch with the table `pc`, as callers run it (the
// eager path: launches, the read-back of the fullest bucket, the host
// doubling chains); best of two after a warm-up
float time_table(MsmEngine &eng, const MsmContext &ctx, const std::vector<MsmJob> &jobs, const PrecompTable &pc,
                 hipStream_t st)
{
    std::vector<RawG1Projective> out(jobs.size());
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        const auto t0 = std::chrono::steady_clock::now();
        eng.run(ctx.points.get(), ctx.npoints, jobs, 0, out.data(), st, &pc);
        const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rep) best = std::min(best, ms);
    }
    return best;
}

}  // namespace

namespace {

// The small table: window kSmallTableC over the points below 2^(pc_c - 1),
// built for autotuned contexts whose main table is wider (a 2^20-point SRS
// serving proofs of k = 10..15 is the case).  At c = 13 an MSM of 2^12
// points has 20 x 2^12 entries in 4096 buckets, against 32 x 2^12 entries
// in 32 x 128 buckets and a 256-bit host chain on plain 8-bit windows.  (At
// c = 12 the top window holds two bits of the scalar: its 2^12 digits fall
// into four buckets, whose chunk trees then set the MSM's latency.)
constexpr uint32_t kSmallTableC = 13;
constexpr uint32_t kMaxTableC = 17;   // widest autotuned or configured table window

void build_small_table(MsmContext &ctx, hipStream_t st)
{
    if (ctx.window_bits || !ctx.pc_nwin || ctx.pc_c <= kSmallTableC) return;
    const size_t rows = std::min<size_t>(ctx.npoints, size_t(1) << (ctx.pc_c - 1));
    if (!build_table(ctx, kSmallTableC, ctx.small_table, st, rows)) return;
    ctx.sm_c = kSmallTableC;
    ctx.sm_nwin = msm_windows(kSmallTableC);
    ctx.sm_n = rows;
}

// The main per-window table.  A config with window_bits set gets that
// window (10..17; the reference's 18 is capped at 17: two key ranges of the
// LDS digit sort).  Otherwise contexts of 2^16 points or more are autotuned
// here, once: a batch of four MSMs of uniform scalars over the whole point
// set is timed with the rule's window and one bit either side (within
// 10..17), each with accumulation chunks of 64 and 32 entries, and the
// fastest table and chunk length are kept (the reference's counterpart is
// its hand-measured per-k table, lib.rs:159-203).  At 17 bits the scalars'
// negation above (r - 1) / 2 saves a window (15 instead of 16) and the
// two-level bucket weights keep its 2^16 buckets' reduction short.  MSM
// results do not depend on the choice.
void build_precomp_main(MsmContext &ctx, hipStream_t st)
{
    const size_t n = ctx.npoints;
    if (n < (size_t(1) << 10)) return;
    const uint32_t wb = ctx.window_bits;
    uint32_t c0 = wb ? std::min<uint32_t>(kMaxTableC, std::max<uint32_t>(10, wb)) : rule_table_window(n);
    // untuned contexts avoid the widths whose top window is degenerate
    // (tuned ones time them: one batch of full-length MSMs may still gain)
    if (!wb && n < (size_t(1) << 16))
        while (!full_top_window(c0)) ++c0;
    if (wb || n < (size_t(1) << 16)) {
        if (build_table(ctx, c0, ctx.table, st)) {
            ctx.pc_c = c0;
            ctx.pc_nwin = msm_windows(c0);
        }
        return;
    }
    DevBuf<Fr> sc(n);
    uint8_t seed[32] = {0};
    seed[0] = 0x6d;   // any fixed seed: uniform scalars
    random_fr_device(sc.get(), n, seed, st);
    // a batch of four MSMs (the batches a prover commits: advice columns,
    // quotient pieces), which is what the chunk length trades against
    const std::vector<MsmJob> jobs(4, MsmJob{sc.get(), 0, n});
    MsmEngine eng;
    float best = 1e30f;
    for (uint32_t c = std::max<uint32_t>(10, c0 - 1); c <= std::min<uint32_t>(kMaxTableC, c0 + 1); ++c) {
        DevBuf<uint32_t> table;
        if (!build_table(ctx, c, table, st)) continue;
        const uint32_t nwin = msm_windows(c);
        const PrecompTable pc{table.get(), n, c, nwin};
        bool keep = false;
        for (uint32_t L : {64u, 32u}) {
            eng.chunk_len = L;
            const float ms = time_table(eng, ctx, jobs, pc, st);
            if (ms < best) {
                best = ms;
                ctx.chunk_len = L;
                ctx.pc_c = c;
                ctx.pc_nwin = nwin;
                keep = true;
            }
        }
        if (keep) ctx.table = std::move(table);
    }
    ctx.tuned_ms = best;
    MH_HIP(hipStreamSynchronize(st));
}

}  // namespace

void build_precomp(MsmContext &ctx, hipStream_t st)
{
    build_precomp_main(ctx, st);
    build_small_table(ctx, st);
}

bool context_table(const MsmContext &ctx, size_t max_n, size_t max_end, PrecompTable &pc)
{
    if (ctx.pc_nwin && max_n >= (size_t(1) << (ctx.pc_c - 1))) {
        pc = PrecompTable{ctx.table.get(), ctx.npoints, ctx.pc_c, ctx.pc_nwin};
        return true;
    }
    if (ctx.sm_nwin && max_end <= ctx.sm_n) {
        pc = PrecompTable{ctx.small_table.get(), ctx.sm_n, ctx.sm_c, ctx.sm_nwin};
        return true;
    }
    return false;
}

// ---------------------------------------------------------------- engine

struct MsmEngine::Buffers {
    DevBuf<uint32_t> cnt, off, cursor, q, co, cmap, ent, scan_tmp, totals;
    DevBuf<uint64_t> tmp;                          // (key, entry) pairs between the placement passes
    DevBuf<uint32_t> ccursor, cst2, cmap2, t2;   // coarse bins: cursors, pass-2 block offsets, map, total
    DevBuf<uint32_t> lc, lo, lmap;   // tree levels: groups per bucket, their offsets, lane -> bucket
    DevBuf<uint32_t> part, part2, bucket, pp, pp2, pj, lines;
    DevBuf<uint32_t> seg_off, seg_len, out_off;
    DevBuf<MsmDesc> descs;
    // pinned host copies of the (window, plane) sums and the totals
    void *h_pj = nullptr;
    size_t h_pj_cap = 0;
    uint32_t *h_totals = nullptr;
    ~Buffers()
    {
        if (h_pj) (void)hipHostFree(h_pj);
        if (h_totals) (void)hipHostFree(h_totals);
    }
};

// What prepare() decided for the next enqueue() / finish().
struct MsmEngine::State {
    struct Level {
        uint64_t base, total_out;
        uint32_t G;
    };
    Geometry g{};
    PlaneLayout pl{};
    const uint32_t *d_points = nullptr;
    uint64_t max_n = 0, nkeys = 0, nent_max = 0, nch_max = 0, npp = 0, nseg = 0;
    // two-pass placement: 2^fshift keys per coarse bin, nbins bins, grid of pass 2
    bool two_pass = false;
    uint32_t fshift = 0, nbins = 0;
    uint64_t fine_blocks = 0;
    uint32_t L = 0, G = 0, tree_levels = 0, G1 = 0;
    bool precomputed = false, ready = false;
    std::vector<Level> levels;
    // two-level weights (precomputed windows of 13 bits or more): line sums
    // first, then the planes over them
    bool lines = false;
    LineLayout ll{};
    uint32_t G1l = 0;
    uint64_t nppl = 0, nsegl = 0;
    std::vector<Level> llevels;
    std::vector<MsmDesc> hd;
    std::vector<uint32_t> seg_off, seg_len, out_off;
    // the layout last uploaded and the device buffers it went to
    std::vector<uint32_t> up_seg_off, up_seg_len, up_out_off;
    const void *up_dev[3] = {nullptr, nullptr, nullptr};
};

MsmEngine::MsmEngine() : buf_(new Buffers), st_(new State) {}
MsmEngine::~MsmEngine() = default;

namespace {

// levels of tile sums of an exclusive scan of n values
std::vector<uint64_t> scan_levels(uint64_t n)
{
    std::vector<uint64_t> sizes;
    uint64_t m = n;
    while (true) {
        sizes.push_back(m);
        if (m <= kScanTile) break;
        m = ceil_div(m, kScanTile);
    }
    return sizes;
}

// scratch elements scan_launch needs for n values
uint64_t scan_tmp_elems(uint64_t n)
{
    const auto sizes = scan_levels(n);
    uint64_t need = 0;
    for (size_t l = 1; l < sizes.size(); ++l) need += 2 * sizes[l];
    return need + 2;
}

// exclusive scan of n u32 values on the device; kernels only (no
// allocation, no read-back), tmp holds scan_tmp_elems(n)
void scan_launch(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *tmp, hipStream_t st)
{
    if (n > kScanTile && n <= kSmallScan) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kSmallScanThreads), 0, st, in, out, tmp, n);
        MH_HIP(hipGetLastError());
        return;
    }
    const auto sizes = scan_levels(n);
    std::vector<uint32_t *> sums(sizes.size(), nullptr), sscan(sizes.size(), nullptr);
    uint64_t o = 0;
    for (size_t l = 1; l < sizes.size(); ++l) {
        sums[l] = tmp + o;
        o += sizes[l];
        sscan[l] = tmp + o;
        o += sizes[l];
    }
    uint32_t *total_dev = tmp + o;
    // up-sweep
    for (size_t l = 0; l < sizes.size(); ++l) {
        const uint32_t *src = l == 0 ? in : sums[l];
        uint32_t *dst = l == 0 ? out : sscan[l];
        uint32_t *next = l + 1 < sizes.size() ? sums[l + 1] : total_dev;
        const uint64_t tiles = ceil_div(sizes[l], kScanTile);
        hipLaunchKernelGGL(k_scan_tiles, dim3((uint32_t)tiles), dim3(kBlock), 0, st, src, dst, next, sizes[l]);
        MH_HIP(hipGetLastError());
    }
    // down-sweep: add the scanned tile sums of level l+1 into level l
    for (size_t l = sizes.size() - 1; l-- > 0;) {
        const uint64_t tiles = ceil_div(sizes[l], kScanTile);
        uint32_t *target = l == 0 ? out : sscan[l];
        hipLaunchKernelGGL(k_scan_add, dim3((uint32_t)tiles), dim3(kBlock), 0, st, target, sscan[l + 1], sizes[l]);
        MH_HIP(hipGetLastError());
    }
}

uint32_t auto_window(uint64_t n)
{
    // measured on MI355X: larger windows cut entries (N * ceil(256/c)) but
    // grow the bucket planes; c around log2(n) - 7 balances the two
    int lg = 0;
    while ((1ull << lg) < n) ++lg;
    int c = lg - 7;
    if (c < 8) c = 8;
    if (c > 16) c = 16;
    return (uint32_t)c;
}

}  // namespace

uint32_t msm_auto_window(uint64_t n) { return auto_window(n); }

void MsmEngine::prepare(const uint32_t *d_points, size_t npoints_total, const std::vector<MsmJob> &jobs,
                        uint32_t window_bits, hipStream_t st, const PrecompTable *pc)
{
    MH_CHECK(!jobs.empty(), (int)hipErrorInvalidValue, "empty MSM batch");
    Buffers &b = *buf_;
    State &s = *st_;
    s.ready = false;
    uint64_t max_n = 0, total_pairs = 0;
    for (const auto &j : jobs) {
        MH_CHECK(j.pt_off + j.n <= npoints_total, (int)hipErrorInvalidValue, "MSM range exceeds the point set");
        MH_CHECK(!pc || j.pt_off + j.n <= pc->n, (int)hipErrorInvalidValue, "MSM range exceeds the table");
        max_n = std::max<uint64_t>(max_n, j.n);
        total_pairs += j.n;
    }
    Geometry &g = s.g;
    s.precomputed = pc != nullptr;
    if (pc) {
        // one bucket set per MSM: window w reads the points [2^(c w)] P_i
        g.c = pc->c;
        g.nwin = pc->nwin;
        g.kw = 1;
        g.pstride = (uint32_t)pc->n;
        d_points = pc->table;
    } else {
        g.c = window_bits ? window_bits : auto_window(max_n);
        g.nwin = msm_windows(g.c);
        g.kw = g.nwin;
        g.pstride = 0;
    }
    MH_CHECK(g.c >= 8 && g.c <= kMaxWindow, (int)hipErrorInvalidValue, "unsupported MSM window width");
    g.B = 1u << (g.c - 1);
    g.M = (uint32_t)jobs.size();
    s.d_points = d_points;
    s.max_n = max_n;
    last_window = g.c;
    // every scalar gives at most nwin nonzero digits (the top window has a
    // spare bit, so no carry leaves it)
    s.nkeys = g.keys();
    s.nent_max = (uint64_t)g.nwin * total_pairs;
    MH_CHECK(s.nent_max < (1ull << 32) && s.nkeys < (1ull << 31), (int)hipErrorInvalidValue, "MSM batch too large");
    // chunk length: enough lanes to fill the GPU, and at least the square
    // root of the mean bucket length, which balances a bucket's serial
    // chain (L mixed additions per chunk, then its ~mean/L chunk partials)
    // on small batches that the fill rule would cut into four-entry chunks
    uint32_t root = 1;
    const uint64_t mean = s.nent_max / std::max<uint64_t>(s.nkeys, 1);
    while ((uint64_t)root * root < mean) ++root;
    s.L = std::max(fill_group(s.nent_max, chunk_len), std::min(root, chunk_len));
    s.nch_max = ceil_div(s.nent_max, s.L) + s.nkeys;
    MH_CHECK(s.nch_max < (1ull << 32), (int)hipErrorInvalidValue, "MSM batch too large");
    // partials per group in the bucket tree.  Sized for the worst case, the
    // tree's levels are launched whether or not a bucket needs them (on
    // uniform scalars the upper levels find no work), so few levels of
    // wide groups beat many narrow ones: small MSMs run their groups on
    // lane quads, and a proof service is bound by launches per proof.
    s.G = 16;
    // fullest possible bucket: every digit of every scalar of the longest
    // MSM in one bucket (one bucket set spans all windows when precomputed)
    const uint64_t bucket_max = pc ? max_n * g.nwin : max_n;
    const uint64_t q_worst = ceil_div(bucket_max, s.L);
    s.tree_levels = 0;   // levels until every bucket holds at most G partials
    for (uint64_t d = s.G; d < q_worst; d *= s.G) ++s.tree_levels;

    b.descs.ensure(jobs.size());
    b.cnt.ensure(s.nkeys);
    b.off.ensure(s.nkeys);
    b.cursor.ensure(s.nkeys);
    b.q.ensure(s.nkeys);
    b.co.ensure(s.nkeys);
    b.scan_tmp.ensure(scan_tmp_elems(s.nkeys));
    b.totals.ensure(4);
    b.ent.ensure(std::max<uint64_t>(s.nent_max, 1));
    s.fshift = 0;
    while ((s.nkeys + (1ull << s.fshift) - 1) >> s.fshift > kCoarseBins) ++s.fshift;
    s.two_pass = (1ull << s.fshift) <= kMaxFine;
    if (s.two_pass) {
        s.nbins = (uint32_t)((s.nkeys + (1ull << s.fshift) - 1) >> s.fshift);
        s.fine_blocks = ceil_div(s.nent_max, kFineChunk) + s.nbins;
        b.tmp.ensure(std::max<uint64_t>(s.nent_max, 1));
        b.ccursor.ensure(s.nbins);
        b.cst2.ensure(s.nbins);
        b.cmap2.ensure(s.fine_blocks);
        b.t2.ensure(1);
    }
    b.cmap.ensure(s.nch_max);
    b.part.ensure(s.nch_max * kXyzz28Words);
    if (s.tree_levels > 0) {
        b.part2.ensure(s.nch_max * kXyzz28Words);
        b.lc.ensure(s.nkeys);
        b.lo.ensure(s.nkeys);
        b.lmap.ensure(s.nch_max / s.G + s.nkeys);
    }
    b.bucket.ensure(s.nkeys * kXyzz28Words);

    // bit planes: S = sum_k 2^k Q_k, Q_k = sum of the buckets x with bit k of
    // x -- or, for a precomputed window of 13 bits or more, of the line sums
    // of the two-level weights (LineLayout)
    const uint64_t nmw = (uint64_t)g.M * g.kw;
    s.lines = g.kw == 1 && g.c >= 13;
    s.seg_off.clear();
    s.seg_len.clear();
    s.out_off.clear();
    // group passes within each segment until one partial per segment is
    // left; returns the level list (descriptors appended to seg_*)
    auto group_levels = [&](std::vector<uint32_t> offs, std::vector<uint32_t> lens, uint64_t in_total) {
        std::vector<State::Level> out;
        const uint64_t nseg = offs.size();
        for (;;) {
            const uint32_t maxlen = *std::max_element(lens.begin(), lens.end());
            const bool last = maxlen <= plane_group2;
            const uint32_t G =
                last ? maxlen : (in_total >= (uint64_t)plane_group2 * kFillLanes / 8 ? plane_group2 : 4);
            State::Level lv{s.seg_off.size(), 0, G};
            for (uint64_t x = 0; x < nseg; ++x) {
                s.seg_off.push_back(offs[x]);
                s.seg_len.push_back(lens[x]);
                s.out_off.push_back((uint32_t)lv.total_out);
                offs[x] = (uint32_t)lv.total_out;
                lens[x] = (lens[x] + G - 1) / G;
                lv.total_out += lens[x];
            }
            out.push_back(lv);
            in_total = lv.total_out;
            if (last) break;
        }
        return out;
    };
    uint64_t pp_need = 0, pp2_need = 1;
    PlaneLayout &pl = s.pl;
    pl = PlaneLayout{};
    pl.c = g.c;
    if (s.lines) {
        LineLayout &ll = s.ll;
        ll.h = g.c / 2;
        ll.H = g.B >> ll.h;
        ll.nlines = ((1u << ll.h) - 1u) + ll.H;
        s.G1l = fill_group((uint64_t)g.M * 2 * g.B, plane_group);
        ll.chunks_r = (ll.H + s.G1l - 1) / s.G1l;
        ll.chunks_c = ((1u << ll.h) + s.G1l - 1) / s.G1l;
        ll.per_msm = ((1u << ll.h) - 1u) * ll.chunks_r + (ll.H - 1) * ll.chunks_c + 1;
        s.nppl = (uint64_t)g.M * ll.per_msm;
        s.nsegl = (uint64_t)g.M * ll.nlines;
        std::vector<uint32_t> offs(s.nsegl), lens(s.nsegl);
        for (uint64_t x = 0; x < s.nsegl; ++x) {
            const uint32_t m = (uint32_t)(x / ll.nlines), line = (uint32_t)(x % ll.nlines);
            const uint32_t nr = (1u << ll.h) - 1u;
            uint32_t first, len;
            if (line < nr) {
                first = line * ll.chunks_r;
                len = ll.chunks_r;
            } else if (line - nr < ll.H - 1) {
                first = nr * ll.chunks_r + (line - nr) * ll.chunks_c;
                len = ll.chunks_c;
            } else {
                first = nr * ll.chunks_r + (ll.H - 1) * ll.chunks_c;
                len = 1;
            }
            offs[x] = m * ll.per_msm + first;
            lens[x] = len;
        }
        s.llevels = group_levels(offs, lens, s.nppl);
        pp_need = std::max(pp_need, s.nppl);
        for (size_t i = 0; i < s.llevels.size(); ++i)
            (i & 1 ? pp_need : pp2_need) = std::max(i & 1 ? pp_need : pp2_need, s.llevels[i].total_out);
        b.lines.ensure(s.nsegl * kXyzz28Words);
        // planes over the lines: k < h over R lines (2^(h-1) members), then
        // over C lines (H / 2 members, one for the top plane)
        pl.h = ll.h;
        pl.nlines = ll.nlines;
        for (uint32_t k = 0; k < g.c; ++k)
            pl.members[k] = k < ll.h ? (1u << (ll.h - 1)) : (k + 1 < g.c ? ll.H / 2 : 1);
        s.G1 = fill_group(nmw * ((uint64_t)g.c * (1u << (ll.h - 1))), plane_group);
    } else {
        s.llevels.clear();
        s.nppl = s.nsegl = 0;
        for (uint32_t k = 0; k < g.c; ++k) pl.members[k] = k + 1 < g.c ? g.B / 2 : 1;
        s.G1 = fill_group(nmw * ((uint64_t)(g.c - 1) * (g.B / 2) + 1), plane_group);
    }
    uint32_t acc = 0;
    for (uint32_t k = 0; k < g.c; ++k) {
        pl.first[k] = acc;
        acc += (pl.members[k] + s.G1 - 1) / s.G1;
    }
    pl.first[g.c] = acc;
    for (uint32_t k = g.c + 1; k <= kMaxWindow; ++k) pl.first[k] = 0xffffffffu;
    pl.per_window = acc;
    s.npp = nmw * pl.per_window;

    // segments (msm, window, plane); the last pass converts to Jacobian
    s.nseg = nmw * g.c;
    {
        std::vector<uint32_t> offs(s.nseg), lens(s.nseg);
        for (uint64_t x = 0; x < s.nseg; ++x) {
            const uint32_t k = (uint32_t)(x % g.c);
            offs[x] = (uint32_t)((x / g.c) * pl.per_window + pl.first[k]);
            lens[x] = pl.first[k + 1] - pl.first[k];
        }
        s.levels = group_levels(offs, lens, s.npp);
    }
    pp_need = std::max(pp_need, s.npp);
    for (size_t i = 0; i < s.levels.size(); ++i)
        (i & 1 ? pp_need : pp2_need) = std::max(i & 1 ? pp_need : pp2_need, s.levels[i].total_out);
    b.pp.ensure(pp_need * kXyzz28Words);
    b.pp2.ensure(pp2_need * kXyzz28Words);
    b.seg_off.ensure(s.seg_off.size());
    b.seg_len.ensure(s.seg_len.size());
    b.out_off.ensure(s.out_off.size());
    b.pj.ensure(s.nseg * 36 + 4);

    // uploads of the descriptors and the plane-pass layout
    s.hd.resize(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i)
        s.hd[i] = MsmDesc{jobs[i].scalars, (uint32_t)jobs[i].pt_off, (uint32_t)jobs[i].n};
    MH_HIP(hipMemcpyAsync(b.descs.get(), s.hd.data(), s.hd.size() * sizeof(MsmDesc), hipMemcpyHostToDevice, st));
    // the plane-pass layout depends on the geometry only: the MSMs of one
    // proof (same size, same engine) upload it once
    const void *dev_now[3] = {b.seg_off.get(), b.seg_len.get(), b.out_off.get()};
    if (s.seg_off != s.up_seg_off || s.seg_len != s.up_seg_len || s.out_off != s.up_out_off ||
        std::memcmp(dev_now, s.up_dev, sizeof(dev_now)) != 0) {
        MH_HIP(hipMemcpyAsync(b.seg_off.get(), s.seg_off.data(), s.seg_off.size() * 4, hipMemcpyHostToDevice, st));
        MH_HIP(hipMemcpyAsync(b.seg_len.get(), s.seg_len.data(), s.seg_len.size() * 4, hipMemcpyHostToDevice, st));
        MH_HIP(hipMemcpyAsync(b.out_off.get(), s.out_off.data(), s.out_off.size() * 4, hipMemcpyHostToDevice, st));
        s.up_seg_off = s.seg_off;
        s.up_seg_len = s.seg_len;
        s.up_out_off = s.out_off;
        std::memcpy(s.up_dev, dev_now, sizeof(dev_now));
    }
    static_assert(sizeof(HJac) == 144, "HJac layout");
    if (b.h_pj_cap < s.nseg) {
        if (b.h_pj) MH_HIP(hipHostFree(b.h_pj));
        b.h_pj = nullptr;
        b.h_pj_cap = 0;
        MH_HIP(hipHostMalloc(&b.h_pj, s.nseg * 144 + 16));   // + the batch totals
        b.h_pj_cap = s.nseg;
    }
    if (!b.h_totals) MH_HIP(hipHostMalloc(reinterpret_cast<void **>(&b.h_totals), 16));
    s.ready = true;
}

void MsmEngine::enqueue_front(hipStream_t st, ProfileScope *prof)
{
    Buffers &b = *buf_;
    State &s = *st_;
    MH_CHECK(s.ready, -1, "MSM engine used before prepare");
    const Geometry &g = s.g;
    const uint64_t nkeys = s.nkeys;
    if (prof) prof->begin("msm_total", st);

    // 1-3: histogram, offsets, scatter
    hipLaunchKernelGGL(k_clear, dim3((uint32_t)std::min<uint64_t>(grid1(nkeys), 4096)), dim3(kBlock), 0, st,
                       b.cnt.get(), nkeys, b.totals.get());
    MH_HIP(hipGetLastError());
    if (s.max_n) launch_digits(false, g.c, s.max_n, st, b.descs.get(), g, b.cnt.get(), nullptr);
    const bool small = nkeys <= kSmallScan;
    if (small) {
        hipLaunchKernelGGL(k_offsets_small, dim3(1), dim3(kSmallScanThreads), 0, st, (const uint32_t *)b.cnt.get(),
                           b.off.get(), b.cursor.get(), nkeys);
        MH_HIP(hipGetLastError());
    } else {
        scan_launch(b.cnt.get(), b.off.get(), nkeys, b.scan_tmp.get(), st);
        hipLaunchKernelGGL(k_copy_u32, dim3((uint32_t)std::min<uint64_t>(grid1(nkeys), 4096)), dim3(kBlock), 0, st,
                           b.cursor.get(), (const uint32_t *)b.off.get(), nkeys);
        MH_HIP(hipGetLastError());
    }
    if (s.max_n && s.two_pass) {
        hipLaunchKernelGGL(k_bin_plan, dim3(1), dim3(kSmallScanThreads), 0, st, (const uint32_t *)b.off.get(),
                           (const uint32_t *)b.cnt.get(), nkeys, s.fshift, s.nbins, b.ccursor.get(), b.cst2.get(),
                           b.cmap2.get(), b.t2.get());
        MH_HIP(hipGetLastError());
        launch_place_coarse(g.c, s.max_n, st, b.descs.get(), g, s.fshift, s.nbins, b.ccursor.get(), b.tmp.get());
        hipLaunchKernelGGL(k_place_fine, dim3((uint32_t)s.fine_blocks), dim3(kPlaceThreads), 0, st,
                           (const uint64_t *)b.tmp.get(), (const uint32_t *)b.off.get(),
                           (const uint32_t *)b.cnt.get(), (const uint32_t *)b.cmap2.get(),
                           (const uint32_t *)b.cst2.get(), (const uint32_t *)b.t2.get(), nkeys, s.fshift,
                           b.cursor.get(), b.ent.get());
        MH_HIP(hipGetLastError());
    } else if (s.max_n) {
        launch_digits(true, g.c, s.max_n, st, b.descs.get(), g, b.cursor.get(), b.ent.get());
    }

    // 4-6: chunked accumulation into buckets
    if (small) {
        hipLaunchKernelGGL(k_chunk_plan_small, dim3(1), dim3(kSmallScanThreads), 0, st, (const uint32_t *)b.cnt.get(),
                           (const uint32_t *)b.off.get(), b.q.get(), b.co.get(), b.cmap.get(), b.totals.get(), nkeys,
                           s.L);
        MH_HIP(hipGetLastError());
    } else {
        hipLaunchKernelGGL(k_chunk_counts, dim3(grid1(nkeys)), dim3(kBlock), 0, st, b.cnt.get(), b.q.get(), nkeys,
                           s.L, b.totals.get());
        MH_HIP(hipGetLastError());
        scan_launch(b.q.get(), b.co.get(), nkeys, b.scan_tmp.get(), st);
        hipLaunchKernelGGL(k_totals, dim3(1), dim3(64), 0, st, b.cnt.get(), b.off.get(), b.q.get(), b.co.get(),
                           nkeys, b.totals.get());
        MH_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_chunk_map, dim3(grid1(nkeys)), dim3(kBlock), 0, st, b.q.get(), b.co.get(), b.cmap.get(),
                           nkeys);
        MH_HIP(hipGetLastError());
    }
    if (prof) prof->begin("msm_accumulate", st);
    if (s.precomputed) {
        if (s.nch_max < quad_below)
            hipLaunchKernelGGL(k_accumulate28<4>, dim3(grid1(4 * s.nch_max)), dim3(kBlock), 0, st, s.d_points,
                               b.ent.get(), b.cnt.get(), b.off.get(), b.q.get(), b.co.get(), b.cmap.get(),
                               b.totals.get(), s.L, b.part.get(), b.bucket.get());
        else
            hipLaunchKernelGGL(k_accumulate28<1>, dim3(grid1(s.nch_max)), dim3(kBlock), 0, st, s.d_points,
                               b.ent.get(), b.cnt.get(), b.off.get(), b.q.get(), b.co.get(), b.cmap.get(),
                               b.totals.get(), s.L, b.part.get(), b.bucket.get());
    } else {
        hipLaunchKernelGGL(k_accumulate, dim3(grid1(s.nch_max)), dim3(kBlock), 0, st, s.d_points, b.ent.get(),
                           b.cnt.get(), b.off.get(), b.q.get(), b.co.get(), b.cmap.get(), b.totals.get(), s.L,
                           b.part.get(), b.bucket.get());
    }
    MH_HIP(hipGetLastError());
    if (prof) prof->end(st);
}

void MsmEngine::enqueue_back(hipStream_t st, uint32_t levels, ProfileScope *prof)
{
    Buffers &b = *buf_;
    State &s = *st_;
    const Geometry &g = s.g;
    const uint64_t nkeys = s.nkeys;
    // lanes per sum of a reduction pass: passes of fewer sums than
    // group16_below run each sum on four quads, of fewer than quad_below on
    // one quad, else on one lane
    const auto width = [this](uint64_t sums) { return sums < group16_below ? 16 : sums < quad_below ? 4 : 1; };
    levels = std::min(levels, s.tree_levels);

    // bucket tree: dense levels, then one sum per bucket
    uint32_t *tin = b.part.get(), *tout = b.part2.get();
    uint64_t div = 1;
    for (uint32_t l = 0; l < levels; ++l) {
        if (nkeys <= kSmallScan) {
            hipLaunchKernelGGL(k_level_plan_small, dim3(1), dim3(kSmallScanThreads), 0, st, (const uint32_t *)b.q.get(),
                               nkeys, div, s.G, b.lo.get(), b.lmap.get(), b.totals.get());
            MH_HIP(hipGetLastError());
        } else {
            hipLaunchKernelGGL(k_level_counts, dim3(grid1(nkeys)), dim3(kBlock), 0, st, b.q.get(), nkeys, div, s.G,
                               b.lc.get());
            MH_HIP(hipGetLastError());
            scan_launch(b.lc.get(), b.lo.get(), nkeys, b.scan_tmp.get(), st);
            hipLaunchKernelGGL(k_level_map, dim3(grid1(nkeys)), dim3(kBlock), 0, st, (const uint32_t *)b.lc.get(),
                               (const uint32_t *)b.lo.get(), nkeys, b.lmap.get(), b.totals.get());
            MH_HIP(hipGetLastError());
        }
        const uint64_t lanes = s.nch_max / (div * s.G) + nkeys;   // bound on the level's groups
        const int w = width(lanes);
        auto *tl = w == 16 ? k_tree_level<16> : w == 4 ? k_tree_level<4> : k_tree_level<1>;
        hipLaunchKernelGGL(tl, dim3(grid1(w * lanes)), dim3(kBlock), 0, st, b.q.get(), b.co.get(), b.lmap.get(),
                           b.lo.get(), b.totals.get(), div, s.G, (const uint32_t *)tin, tout);
        MH_HIP(hipGetLastError());
        div *= s.G;
        std::swap(tin, tout);
    }
    {
        const int w = width(nkeys);
        auto *tf = w == 16 ? k_tree_final<16> : w == 4 ? k_tree_final<4> : k_tree_final<1>;
        hipLaunchKernelGGL(tf, dim3(grid1(w * nkeys)), dim3(kBlock), 0, st, b.q.get(), b.co.get(), nkeys, s.G, levels,
                           (const uint32_t *)b.part.get(), (const uint32_t *)b.part2.get(), b.bucket.get());
        MH_HIP(hipGetLastError());
    }

    // group passes of a segment set down to one value per segment (the
    // last pass writes `final`, as Jacobian when `to_jac`)
    auto group_passes = [&](const std::vector<State::Level> &lvls, uint64_t nseg, uint32_t *final, int to_jac) {
        uint32_t *lin = b.pp.get(), *lout = b.pp2.get();
        for (size_t i = 0; i < lvls.size(); ++i) {
            const State::Level &lv = lvls[i];
            const bool last = i + 1 == lvls.size();
            const int w = width(lv.total_out);
            auto *gs = w == 16 ? k_group_sums<16> : w == 4 ? k_group_sums<4> : k_group_sums<1>;
            hipLaunchKernelGGL(gs, dim3(grid1(w * lv.total_out)), dim3(kBlock), 0, st, lin, nseg,
                               b.seg_off.get() + lv.base, b.seg_len.get() + lv.base, b.out_off.get() + lv.base, lv.G,
                               lv.total_out, last ? final : lout, last ? to_jac : 0,
                               last && to_jac ? (const uint32_t *)b.totals.get() : nullptr);
            MH_HIP(hipGetLastError());
            std::swap(lin, lout);
        }
    };

    // 7: two-level weights: line sums of the buckets, then planes over them
    const uint32_t *plane_src = b.bucket.get();
    if (s.lines) {
        const int w = width(s.nppl);
        auto *lp = w == 16 ? k_line_partials<16> : w == 4 ? k_line_partials<4> : k_line_partials<1>;
        hipLaunchKernelGGL(lp, dim3(grid1(w * s.nppl)), dim3(kBlock), 0, st, b.bucket.get(), b.cnt.get(), g, s.ll,
                           s.G1l, s.nppl, b.pp.get());
        MH_HIP(hipGetLastError());
        group_passes(s.llevels, s.nsegl, b.lines.get(), 0);
        plane_src = b.lines.get();
    }
    // bit-plane partial sums and group passes down to one per segment
    {
        const int w = width(s.npp);
        auto *pp = w == 16 ? k_plane_partials<16> : w == 4 ? k_plane_partials<4> : k_plane_partials<1>;
        hipLaunchKernelGGL(pp, dim3(grid1(w * s.npp)), dim3(kBlock), 0, st, plane_src, b.cnt.get(), g, s.pl, s.G1,
                           s.npp, b.pp.get());
        MH_HIP(hipGetLastError());
    }
    group_passes(s.levels, s.nseg, b.pj.get(), 1);
    // the (window, plane) sums and, after them, the batch totals
    MH_HIP(hipMemcpyAsync(b.h_pj, b.pj.get(), s.nseg * 144 + 16, hipMemcpyDeviceToHost, st));
    if (prof) prof->end(st);
}

void MsmEngine::enqueue(hipStream_t st, ProfileScope *prof)
{
    enqueue_front(st, prof);
    enqueue_back(st, st_->tree_levels, prof);
}

void MsmEngine::finish(RawG1Projective *out)
{
    const Buffers &b = *buf_;
    const State &s = *st_;
    MH_CHECK(s.ready, -1, "MSM engine used before prepare");
    const Geometry &g = s.g;
    last_entries = static_cast<const uint32_t *>(b.h_pj)[s.nseg * 36];
    const HJac *hj = static_cast<const HJac *>(b.h_pj);
    // 8: host: one doubling chain per MSM over its (window, plane) sums,
    // exponent e = c*w + k
    auto chain = [&](uint32_t m) {
        const uint32_t E = g.kw * g.c;
        HJac accj = HJac::identity();
        for (uint32_t e = E; e-- > 0;) {
            accj = accj.dbl();
            uint32_t w = e / g.c, k = e % g.c;
            accj = accj + hj[((uint64_t)m * g.kw + w) * g.c + k];
        }
        std::memcpy(&out[m], &accj, sizeof(RawG1Projective));
    };
    const uint32_t nthreads = std::min<uint32_t>(g.M, 16);
    if (nthreads <= 1) {
        for (uint32_t m = 0; m < g.M; ++m) chain(m);
    } else {
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < nthreads; ++t)
            th.emplace_back([&, t]() {
                for (uint32_t m = t; m < g.M; m += nthreads) chain(m);
            });
        for (auto &x : th) x.join();
    }
}

namespace {
// test hook: a cap on the bucket-tree levels of the calling thread's eager
// MSMs (the last pass then sums deeper buckets' partials itself)
thread_local uint32_t t_tree_level_cap = ~0u;
}  // namespace

void MsmEngine::run(const uint32_t *d_points, size_t npoints_total, const std::vector<MsmJob> &jobs,
                    uint32_t window_bits, RawG1Projective *out, hipStream_t st, const PrecompTable *pc)
{
    if (jobs.empty()) return;
    prepare(d_points, npoints_total, jobs, window_bits, st, pc);
    ProfileScope prof;
    ProfileScope *pr = profiling() ? &prof : nullptr;
    enqueue_front(st, pr);
    // one read-back (the fullest bucket's chunk count) so that only the
    // tree levels this batch needs are launched
    Buffers &b = *buf_;
    MH_HIP(hipMemcpyAsync(b.h_totals, b.totals.get(), 16, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    const State &s = *st_;
    uint32_t levels = 0;
    for (uint32_t rem = b.h_totals[3]; rem > s.G; rem = (rem + s.G - 1) / s.G) ++levels;
    enqueue_back(st, std::min(levels, t_tree_level_cap), pr);
    MH_HIP(hipStreamSynchronize(st));
    prof.flush();
    finish(out);
}

}  // namespace mh

// ---------------------------------------------------------------- C ABI

using namespace mh;

namespace {

}  // namespace

RawMsmConfig mh::default_msm_config(size_t npoints)
{
    uint32_t k = 0;
    size_t cap = npoints > 1 ? npoints - 1 : 0;
    while (cap) {
        ++k;
        cap >>= 1;
    }
    uint32_t wb = k == 0 ? 0 : 10;
    if (k >= 16 && k <= 18) wb = 12;
    if (k > 18) wb = 0;
    return RawMsmConfig{k, 100000, 32768, wb, 3, 2, 1, 4, 0, (uint32_t)(k >= 14 && k <= 18 ? 1 : 0)};
}

std::unique_ptr<MsmContext> mh::make_context(size_t npoints)
{
    auto ctx = std::make_unique<MsmContext>();
    ctx->lane = current_lane();
    ctx->config = default_msm_config(npoints);
    ctx->npoints = npoints;
    ctx->points.alloc(std::max<size_t>(npoints, 1) * 24);
    return ctx;
}

namespace {

// bls12_381_g1_msm.cu:85-110
bool valid_config(const RawMsmConfig &c)
{
    return c.domain_k <= 31 && c.min_msm_scalars >= 1 && c.unfiltered_min_scalars >= 1 &&
           (c.window_bits == 0 || (c.window_bits >= 10 && c.window_bits <= 18)) && c.sort_grid_divisor >= 1 &&
           c.sort_grid_divisor <= 8 && c.sort_window_lanes >= 1 && c.sort_window_lanes <= 4 &&
           (c.integration_blocks == 1 || c.integration_blocks == 2 || c.integration_blocks == 4 ||
            c.integration_blocks == 8 || c.integration_blocks == 16) &&
           c.batch_lanes >= 1 && c.batch_lanes <= 16 && c.reuse_event <= 1 && c.device_reduce <= 3;
}

// Plain windows follow the window_bits of a caller's config when it is set
// (0 lets the engine pick its width from the MSM size).
uint32_t engine_window(const MsmContext &ctx) { return ctx.window_bits; }

constexpr size_t kMaxBatchPairs = size_t(1) << 26;   // bounds entry buffers (~ 20 windows x 4 B x this)

void run_batched(MsmContext &ctx, MsmEngine &engine, const std::vector<MsmJob> &jobs, RawG1Projective *out,
                 hipStream_t st)
{
    size_t i = 0;
    while (i < jobs.size()) {
        size_t j = i, pairs = 0;
        while (j < jobs.size() && (j == i || pairs + jobs[j].n <= kMaxBatchPairs)) {
            pairs += jobs[j].n;
            ++j;
        }
        std::vector<MsmJob> part(jobs.begin() + i, jobs.begin() + j);
        size_t part_max = 0, part_end = 0;
        for (const auto &pj : part) {
            part_max = std::max<size_t>(part_max, pj.n);
            part_end = std::max<size_t>(part_end, pj.pt_off + pj.n);
        }
        // the main table's one bucket set has 2^(c-1) buckets; MSMs much
        // shorter than that (a small circuit over a large SRS) take the small
        // table when it covers them, else plain windows sized for their length
        engine.chunk_len = ctx.chunk_len;
        PrecompTable pc;
        if (context_table(ctx, part_max, part_end, pc)) {
            engine.run(ctx.points.get(), ctx.npoints, part, 0, out + i, st, &pc);
        } else {
            engine.run(ctx.points.get(), ctx.npoints, part, engine_window(ctx), out + i, st);
        }
        i = j;
    }
}

RawG1Projective identity_out()
{
    RawG1Projective r;
    HJac id = HJac::identity();
    std::memcpy(&r, &id, sizeof(r));
    return r;
}

}  // namespace

namespace {

// Resident contexts precompute per-window point copies (worth it because a
// context serves many MSMs); the one-shot entry points do not.
RustError create_context(void **out, const RawG1Affine *points, size_t npoints, RawMsmConfig config,
                         bool precompute, bool caller_config)
{
    if (!out) return make_error((int)hipErrorInvalidValue, "out is null");
    *out = nullptr;
    if (!valid_config(config)) return make_error((int)hipErrorInvalidValue, "invalid BLS12-381 MSM config");
    if (npoints && !points) return make_error((int)hipErrorInvalidValue, "points is null");
    if (npoints >= (size_t(1) << 31)) return make_error((int)hipErrorInvalidValue, "too many MSM points");
    MH_TRY
    auto ctx = make_context(npoints);
    ctx->config = config;
    if (caller_config) ctx->window_bits = config.window_bits;
    hipStream_t st = cur_stream();
    if (npoints) MH_HIP(hipMemcpyAsync(ctx->points.get(), points, npoints * 96, hipMemcpyHostToDevice, st));
    if (precompute) build_precomp(*ctx, st);
    MH_HIP(hipStreamSynchronize(st));
    *out = ctx.release();
    return ok();
    MH_CATCH
}

}  // namespace

extern "C" RustError midnight_sppark_bls12_381_g1_msm_context_create_with_config(void **out,
                                                                                 const RawG1Affine *points,
                                                                                 size_t npoints, RawMsmConfig config)
{
    return create_context(out, points, npoints, config, true, true);
}

extern "C" RustError midnight_sppark_bls12_381_g1_msm_context_create(void **out, const RawG1Affine *points,
                                                                     size_t npoints)
{
    return create_context(out, points, npoints, default_msm_config(npoints), true, false);
}

// The table window, accumulation chunk bound and autotuned MSM time (ms; 0
// when the context was not tuned) of a context.
extern "C" RustError midnight_hip_g1_msm_context_tuning(const void *context, uint32_t *window_bits,
                                                        uint32_t *chunk_len, float *tuned_ms)
{
    if (!context) return make_error((int)hipErrorInvalidValue, "context is null");
    const MsmContext *ctx = static_cast<const MsmContext *>(context);
    if (window_bits) *window_bits = ctx->pc_nwin ? ctx->pc_c : ctx->window_bits;
    if (chunk_len) *chunk_len = ctx->chunk_len;
    if (tuned_ms) *tuned_ms = ctx->tuned_ms;
    return ok();
}

extern "C" void midnight_sppark_bls12_381_g1_msm_context_free(void *context)
{
    MsmContext *ctx = static_cast<MsmContext *>(context);
    if (!ctx) return;
    try {
        (void)hipStreamSynchronize(lane_stream(ctx->lane));
    } catch (...) {
    }
    delete ctx;
}

extern "C" RustError midnight_sppark_bls12_381_g1_msm_context_batch(RawG1Projective *out, void *context,
                                                                    const RawFr *scalars, size_t npoints, size_t nmsms)
{
    if (!context) return make_error(-1, "null BLS12-381 G1 MSM context");
    MsmContext *ctx = static_cast<MsmContext *>(context);
    if (nmsms == 0) return ok();
    if (!out || (npoints && !scalars)) return make_error((int)hipErrorInvalidValue, "invalid MSM arguments");
    if (npoints > ctx->npoints) return make_error((int)hipErrorInvalidValue, "MSM length exceeds the context");
    MH_TRY
    MsmContext::LaneEngine &le = ctx->lane_engine(current_lane());
    std::lock_guard<std::mutex> g(le.mu);
    hipStream_t st = cur_stream();
    if (npoints == 0) {
        for (size_t i = 0; i < nmsms; ++i) out[i] = identity_out();
        return ok();
    }
    DevBuf<Fr> d_s(npoints * nmsms);
    MH_HIP(hipMemcpyAsync(d_s.get(), scalars, npoints * nmsms * 32, hipMemcpyHostToDevice, st));
    std::vector<MsmJob> jobs(nmsms);
    for (size_t i = 0; i < nmsms; ++i) jobs[i] = MsmJob{d_s.get() + i * npoints, 0, npoints};
    run_batched(*ctx, le.engine, jobs, out, st);
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_g1_msm_context_device_batch(RawG1Projective *out, void *context,
                                                                           const void *device_vec,
                                                                           size_t scalar_offset, size_t npoints,
                                                                           size_t nmsms)
{
    if (!context || !device_vec || (nmsms && !out))
        return make_error(-1, "invalid resident BLS12-381 G1 MSM arguments");
    MsmContext *ctx = static_cast<MsmContext *>(context);
    const DeviceVec *dv = static_cast<const DeviceVec *>(device_vec);
    if (npoints && nmsms > (SIZE_MAX - scalar_offset) / npoints)
        return make_error((int)hipErrorInvalidValue, "resident MSM range overflow");
    if (scalar_offset + npoints * nmsms > dv->len)
        return make_error((int)hipErrorInvalidValue, "resident MSM range exceeds device vector");
    if (npoints > ctx->npoints) return make_error((int)hipErrorInvalidValue, "MSM length exceeds the context");
    if (nmsms == 0) return ok();
    MH_TRY
    MsmContext::LaneEngine &le = ctx->lane_engine(current_lane());
    std::lock_guard<std::mutex> g(le.mu);
    hipStream_t st = cur_stream();
    if (npoints == 0) {
        for (size_t i = 0; i < nmsms; ++i) out[i] = identity_out();
        return ok();
    }
    std::vector<MsmJob> jobs(nmsms);
    for (size_t i = 0; i < nmsms; ++i) jobs[i] = MsmJob{dv->values.get() + scalar_offset + i * npoints, 0, npoints};
    run_batched(*ctx, le.engine, jobs, out, st);
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_g1_msm_context_device_ragged(RawG1Projective *out, void *context,
                                                                            const void *device_vec,
                                                                            const RawMsmSlice *slices, size_t count)
{
    if (!context || !device_vec || (count && (!out || !slices)))
        return make_error(-1, "invalid ragged resident BLS12-381 G1 MSM arguments");
    if (count == 0) return ok();
    MsmContext *ctx = static_cast<MsmContext *>(context);
    const DeviceVec *dv = static_cast<const DeviceVec *>(device_vec);
    for (size_t i = 0; i < count; ++i) {
        const RawMsmSlice &s = slices[i];
        if (s.point_offset > ctx->npoints || s.len > ctx->npoints - s.point_offset || s.scalar_offset > dv->len ||
            s.len > dv->len - s.scalar_offset || s.nonzero_len > s.len)
            return make_error((int)hipErrorInvalidValue, "ragged resident MSM slice exceeds its source");
    }
    MH_TRY
    MsmContext::LaneEngine &le = ctx->lane_engine(current_lane());
    std::lock_guard<std::mutex> g(le.mu);
    hipStream_t st = cur_stream();
    // zero scalars produce no digits, so the nonzero prefix needs no compaction
    std::vector<MsmJob> jobs;
    std::vector<size_t> where;
    for (size_t i = 0; i < count; ++i) {
        if (slices[i].nonzero_len == 0 || slices[i].len == 0) {
            out[i] = identity_out();
            continue;
        }
        jobs.push_back(MsmJob{dv->values.get() + slices[i].scalar_offset, slices[i].point_offset, slices[i].len});
        where.push_back(i);
    }
    std::vector<RawG1Projective> res(jobs.size());
    run_batched(*ctx, le.engine, jobs, res.data(), st);
    for (size_t i = 0; i < jobs.size(); ++i) out[where[i]] = res[i];
    return ok();
    MH_CATCH
}

// ------------------------------------------------------------- test hooks
//
// bls12_381_g1_msm.cu:6379-6424 (feature "test-hooks"): hold a producer back
// on its lane's stream with a host callback that sleeps, without blocking
// other streams, so a test observes the ordering of producer and consumer
// rather than a lucky schedule.  Every kernel still runs to completion.
namespace {

thread_local uint32_t t_mask_delay_ms = 0;

void sleep_ms_callback(void *ms)
{
    std::this_thread::sleep_for(std::chrono::milliseconds(reinterpret_cast<uintptr_t>(ms)));
}

void delay_stream(hipStream_t st, uint32_t ms)
{
    MH_HIP(hipLaunchHostFunc(st, sleep_ms_callback, reinterpret_cast<void *>(static_cast<uintptr_t>(ms))));
}

}  // namespace

extern "C" void midnight_sppark_test_mask_delay(uint32_t delay_ms) { t_mask_delay_ms = delay_ms; }

extern "C" void midnight_hip_test_msm_tree_levels(uint32_t max_levels) { t_tree_level_cap = max_levels; }

// scale a device vector in place behind a delay of delay_ms on the lane's
// stream; returns at once (the scaling is still pending)
extern "C" RustError midnight_sppark_test_delayed_scale(void *raw_device, RawFr scale, uint32_t delay_ms)
{
    if (!raw_device) return make_error((int)hipErrorInvalidValue, "device vector is null");
    auto *dv = static_cast<DeviceVec *>(raw_device);
    MH_TRY
    hipStream_t st = cur_stream();
    MH_HIP(hipStreamSynchronize(st));
    delay_stream(st, delay_ms);
    fr_mul_scalar(dv->values.get(), to_fr(scale), dv->len, st);
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_g1_msm_context_device_scrambled_batch(
    RawG1Projective *out, void *context, const void *device_vec, size_t scalar_offset, size_t npoints, size_t nmsms,
    const void *mask_device_vec)
{
    if (!context || !device_vec || !mask_device_vec || (nmsms && !out))
        return make_error(-1, "invalid scrambled resident BLS12-381 G1 MSM arguments");
    MsmContext *ctx = static_cast<MsmContext *>(context);
    const DeviceVec *dv = static_cast<const DeviceVec *>(device_vec);
    const DeviceVec *mask = static_cast<const DeviceVec *>(mask_device_vec);
    if (mask->len != npoints) return make_error((int)hipErrorInvalidValue, "resident MSM scramble mask length mismatch");
    if (npoints && nmsms > (SIZE_MAX - scalar_offset) / npoints)
        return make_error((int)hipErrorInvalidValue, "scrambled resident MSM range overflow");
    const size_t count = npoints * nmsms;
    if (scalar_offset + count > dv->len)
        return make_error((int)hipErrorInvalidValue, "scrambled resident MSM range exceeds device vector");
    if (npoints > ctx->npoints) return make_error((int)hipErrorInvalidValue, "MSM length exceeds the context");
    if (count == 0) return ok();
    MH_TRY
    MsmContext::LaneEngine &le = ctx->lane_engine(current_lane());
    std::lock_guard<std::mutex> g(le.mu);
    hipStream_t st = cur_stream();
    // returns MSM(P, s + mask); the caller subtracts MSM(P, mask) (lib.rs:3349-3353)
    DevBuf<Fr> masked(count);
    if (t_mask_delay_ms) {
        // test hook: the masked scalars start as zeros and their producer is
        // held back, so an MSM that read them early would be wrong
        Fr zero;
        std::memset(zero.v, 0, sizeof(zero.v));
        fr_fill(masked.get(), zero, count, st);
        MH_HIP(hipStreamSynchronize(st));
        delay_stream(st, t_mask_delay_ms);
    }
    fr_add_mask(masked.get(), dv->values.get() + scalar_offset, mask->values.get(), npoints, count, st);
    std::vector<MsmJob> jobs(nmsms);
    for (size_t i = 0; i < nmsms; ++i) jobs[i] = MsmJob{masked.get() + i * npoints, 0, npoints};
    run_batched(*ctx, le.engine, jobs, out, st);
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_hip_g1_msm_context_device_multi(RawG1Projective *out, void *context,
                                                              const void *device_vec, const size_t *scalar_offsets,
                                                              const size_t *sizes, size_t count)
{
    if (!context || !device_vec || (count && (!out || !scalar_offsets || !sizes)))
        return make_error(-1, "invalid batched MSM arguments");
    MsmContext *ctx = static_cast<MsmContext *>(context);
    const DeviceVec *dv = static_cast<const DeviceVec *>(device_vec);
    for (size_t i = 0; i < count; ++i)
        if (sizes[i] > ctx->npoints || scalar_offsets[i] > dv->len || sizes[i] > dv->len - scalar_offsets[i])
            return make_error((int)hipErrorInvalidValue, "batched MSM slice exceeds its source");
    if (count == 0) return ok();
    MH_TRY
    MsmContext::LaneEngine &le = ctx->lane_engine(current_lane());
    std::lock_guard<std::mutex> g(le.mu);
    hipStream_t st = cur_stream();
    std::vector<MsmJob> jobs;
    std::vector<size_t> where;
    for (size_t i = 0; i < count; ++i) {
        if (sizes[i] == 0) {
            out[i] = identity_out();
            continue;
        }
        jobs.push_back(MsmJob{dv->values.get() + scalar_offsets[i], 0, sizes[i]});
        where.push_back(i);
    }
    std::vector<RawG1Projective> res(jobs.size());
    run_batched(*ctx, le.engine, jobs, res.data(), st);
    for (size_t i = 0; i < jobs.size(); ++i) out[where[i]] = res[i];
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_g1_msm_batch(RawG1Projective *out, const RawG1Affine *points,
                                                            const RawFr *scalars, size_t npoints, size_t nmsms)
{
    void *h = nullptr;
    RustError e = create_context(&h, points, npoints, default_msm_config(npoints), false, false);
    if (e.code) return e;
    e = midnight_sppark_bls12_381_g1_msm_context_batch(out, h, scalars, npoints, nmsms);
    midnight_sppark_bls12_381_g1_msm_context_free(h);
    return e;
}

extern "C" RustError midnight_sppark_bls12_381_g1_msm(RawG1Projective *out, const RawG1Affine *points,
                                                      const RawFr *scalars, size_t len)
{
    return midnight_sppark_bls12_381_g1_msm_batch(out, points, scalars, len, 1);
}

extern "C" RustError midnight_hip_g1_msm_context_last_stats(const void *context, uint64_t *entries,
                                                           uint32_t *window_bits)
{
    if (!context) return make_error((int)hipErrorInvalidValue, "null context");
    const MsmContext *ctx = static_cast<const MsmContext *>(context);
    MsmContext::LaneEngine &le = const_cast<MsmContext *>(ctx)->lane_engine(current_lane());
    if (entries) *entries = le.engine.last_entries;
    if (window_bits) *window_bits = le.engine.last_window;
    return ok();
}
