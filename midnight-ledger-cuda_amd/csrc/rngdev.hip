// Field elements drawn from ChaCha20Rng streams on the device: the vanishing
// argument's random polynomial (vanishing/prover.rs:41-75), whose rows are
// split into one chunk per prover thread as utils/arithmetic.rs:138-190
// parallelize splits them, chunk c holding Fq::random values of
// ChaCha20Rng::from_seed(seed_c); bit-identical to drawing them one by one
// on the host.
//
// A fresh ChaCha20Rng only ever serves whole u64 pairs here, so candidate j
// (four next_u64) is words [8j, 8j + 8) of the keystream, i.e. one half of
// ChaCha block j / 2 (64-bit block counter, stream 0).  The top bit is
// cleared and the candidate is accepted when below r (about 90.6 % are);
// output i is the i-th accepted candidate.  Every lane makes four blocks
// (eight candidates); a count pass, a scan of the block counts and a
// placement pass (recomputing the blocks, which is cheaper than storing
// them) put the accepted values in order.  All streams of a call share the
// three launches.
#include <cstring>
#include <vector>

#include <memory>

#include "devvec.hpp"
#include "fields.hpp"
#include "host_fr.hpp"
#include "rngdev.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

constexpr int kBlock = 256;
constexpr int kCandPerLane = 8;
constexpr size_t kCandPerBlock = (size_t)kBlock * kCandPerLane;

struct ChaKey {
    uint32_t k[8];
};

__device__ __forceinline__ uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ __forceinline__ void chacha_block(const ChaKey &key, uint64_t counter, uint32_t out[16])
{
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key.k[0], key.k[1], key.k[2], key.k[3],
                      key.k[4],    key.k[5],    key.k[6],    key.k[7],    (uint32_t)counter,
                      (uint32_t)(counter >> 32), 0u, 0u};
    uint32_t x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = s[i];
#define QR(a, b, c, d)                                                      \
    x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 16);                      \
    x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 12);                      \
    x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 8);                       \
    x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 7);
    for (int r = 0; r < 10; ++r) {
        QR(0, 4, 8, 12) QR(1, 5, 9, 13) QR(2, 6, 10, 14) QR(3, 7, 11, 15)
        QR(0, 5, 10, 15) QR(1, 6, 11, 12) QR(2, 7, 8, 13) QR(3, 4, 9, 14)
    }
#undef QR
#pragma unroll
    for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

// candidate from 8 keystream words, top bit cleared; true when below r
__device__ __forceinline__ bool candidate(const uint32_t *w, Fr &c)
{
#pragma unroll
    for (int i = 0; i < 8; ++i) c.v[i] = w[i];
    c.v[7] &= 0x7fffffffu;
    for (int i = 7; i >= 0; --i) {
        const uint32_t p = FrCfg::p(i);
        if (c.v[i] != p) return c.v[i] < p;
    }
    return false;   // equal to r
}

// lane's 8 candidates -> accepted flags (bit i) and, optionally, values
__device__ __forceinline__ uint32_t lane_candidates(const ChaKey &key, size_t lane, Fr *vals)
{
    uint32_t flags = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        uint32_t w[16];
        chacha_block(key, (uint64_t)lane * 4 + b, w);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            Fr c;
            if (candidate(w + 8 * h, c)) flags |= 1u << (2 * b + h);
            if (vals) vals[2 * b + h] = c;
        }
    }
    return flags;
}

// A set of independent streams: segment s fills out[offset_s, offset_s +
// len_s) from ChaCha20Rng::from_seed(seed_s).  Candidate blocks are numbered
// across the set; segment s owns blocks [first_s, first_{s+1}).
struct Segments {
    const ChaKey *keys;
    const uint32_t *first_block;   // count + 1 entries
    const size_t *offset;
    const size_t *len;
    uint32_t count;
};

__device__ __forceinline__ uint32_t segment_of(const Segments &sg, uint32_t block)
{
    uint32_t lo = 0, hi = sg.count - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (sg.first_block[mid] <= block) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void k_rng_count(Segments sg, uint32_t *block_counts)
{
    __shared__ uint32_t sum;
    if (threadIdx.x == 0) sum = 0;
    __syncthreads();
    const uint32_t s = segment_of(sg, blockIdx.x);
    const size_t lane = (blockIdx.x - sg.first_block[s]) * (size_t)kBlock + threadIdx.x;
    atomicAdd(&sum, (uint32_t)__popc(lane_candidates(sg.keys[s], lane, nullptr)));
    __syncthreads();
    if (threadIdx.x == 0) block_counts[blockIdx.x] = sum;
}

// exclusive scan of the block counts within each segment, one lane (a few
// thousand blocks); seg_total[s] = accepted candidates of segment s
__global__ void k_rng_scan(Segments sg, uint32_t *counts, uint32_t *seg_total)
{
    for (uint32_t s = 0; s < sg.count; ++s) {
        uint32_t run = 0;
        for (uint32_t b = sg.first_block[s]; b < sg.first_block[s + 1]; ++b) {
            const uint32_t c = counts[b];
            counts[b] = run;
            run += c;
        }
        seg_total[s] = run;
    }
}

__global__ __launch_bounds__(kBlock) void k_rng_place(Segments sg, const uint32_t *block_offsets, Fr *out, Fr r2)
{
    __shared__ uint32_t sh[2][kBlock];
    const uint32_t s = segment_of(sg, blockIdx.x);
    const size_t lane = (blockIdx.x - sg.first_block[s]) * (size_t)kBlock + threadIdx.x;
    Fr vals[kCandPerLane];
    const uint32_t flags = lane_candidates(sg.keys[s], lane, vals);
    const int t = threadIdx.x;
    sh[0][t] = (uint32_t)__popc(flags);
    __syncthreads();
    int b = 0;
    for (int off = 1; off < kBlock; off <<= 1) {
        uint32_t x = sh[b][t];
        if (t >= off) x += sh[b][t - off];
        sh[b ^ 1][t] = x;
        __syncthreads();
        b ^= 1;
    }
    const size_t n = sg.len[s];
    Fr *dst = out + sg.offset[s];
    size_t pos = block_offsets[blockIdx.x] + (t ? sh[b][t - 1] : 0u);
#pragma unroll
    for (int i = 0; i < kCandPerLane; ++i) {
        if (!((flags >> i) & 1)) continue;
        if (pos < n) store_fr(dst + pos, vals[i] * r2);   // canonical -> Montgomery
        ++pos;
    }
}

ChaKey key_of(const uint8_t seed[32])
{
    ChaKey key;
    for (int i = 0; i < 8; ++i)
        key.k[i] = (uint32_t)seed[4 * i] | (uint32_t)seed[4 * i + 1] << 8 | (uint32_t)seed[4 * i + 2] << 16 |
                   (uint32_t)seed[4 * i + 3] << 24;
    return key;
}

}  // namespace

void random_fr_streams_device(Fr *out, const size_t *offsets, const size_t *lens, const uint8_t *seeds,
                              size_t count, hipStream_t st)
{
    if (count == 0) return;
    MH_CHECK(count < (1u << 20), (int)hipErrorInvalidValue, "too many random streams");
    std::vector<ChaKey> keys(count);
    std::vector<uint32_t> first(count + 1);
    size_t nblocks = 0;
    for (size_t s = 0; s < count; ++s) {
        keys[s] = key_of(seeds + 32 * s);
        first[s] = (uint32_t)nblocks;
        // about 1.104 candidates per value; a wide margin (tens of standard
        // deviations at any length) and a check of the count
        const size_t want = lens[s] + lens[s] / 8 + 4096;
        nblocks += lens[s] ? ceil_div(want, kCandPerBlock) : 0;
        MH_CHECK(nblocks < (1u << 31), (int)hipErrorInvalidValue, "too many random values");
    }
    first[count] = (uint32_t)nblocks;
    if (nblocks == 0) return;
    // one upload: keys | first blocks | offsets | lengths | block counts | totals
    const size_t key_b = count * sizeof(ChaKey), first_b = (count + 1) * 4, size_b = count * sizeof(size_t);
    auto al = [](size_t b) { return (b + 15) & ~size_t(15); };
    const size_t o_first = al(key_b), o_off = o_first + al(first_b), o_len = o_off + al(size_b),
                 o_counts = o_len + al(size_b), o_tot = o_counts + al((nblocks + 1) * 4),
                 bytes = o_tot + al(count * 4);
    std::vector<uint8_t> host(o_counts);
    std::memcpy(host.data(), keys.data(), key_b);
    std::memcpy(host.data() + o_first, first.data(), first_b);
    std::memcpy(host.data() + o_off, offsets, size_b);
    std::memcpy(host.data() + o_len, lens, size_b);
    DevBuf<uint8_t> buf(bytes);
    uint8_t *d = buf.get();
    MH_HIP(hipMemcpyAsync(d, host.data(), o_counts, hipMemcpyHostToDevice, st));
    Segments sg;
    sg.keys = reinterpret_cast<const ChaKey *>(d);
    sg.first_block = reinterpret_cast<const uint32_t *>(d + o_first);
    sg.offset = reinterpret_cast<const size_t *>(d + o_off);
    sg.len = reinterpret_cast<const size_t *>(d + o_len);
    sg.count = (uint32_t)count;
    uint32_t *counts = reinterpret_cast<uint32_t *>(d + o_counts);
    uint32_t *totals = reinterpret_cast<uint32_t *>(d + o_tot);
    hipLaunchKernelGGL(k_rng_count, dim3((unsigned)nblocks), dim3(kBlock), 0, st, sg, counts);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rng_scan, dim3(1), dim3(1), 0, st, sg, counts, totals);
    MH_HIP(hipGetLastError());
    std::vector<uint32_t> tot(count);
    MH_HIP(hipMemcpyAsync(tot.data(), totals, count * 4, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    for (size_t s = 0; s < count; ++s) MH_CHECK(tot[s] >= lens[s], -1, "ChaCha20 candidates exhausted");
    Fr r2;
    std::memcpy(r2.v, HostFr::R2, 32);
    hipLaunchKernelGGL(k_rng_place, dim3((unsigned)nblocks), dim3(kBlock), 0, st, sg, (const uint32_t *)counts, out,
                       r2);
    MH_HIP(hipGetLastError());
}

void random_fr_device(Fr *out, size_t n, const uint8_t seed[32], hipStream_t st)
{
    const size_t off = 0;
    random_fr_streams_device(out, &off, &n, seed, 1, st);
}

void parallelize_chunks(size_t n, size_t threads, std::vector<size_t> &offsets, std::vector<size_t> &lens)
{
    offsets.clear();
    lens.clear();
    if (threads == 0) threads = 1;
    const size_t base = n / threads, cutoff = n % threads, split = cutoff * (base + 1);
    for (size_t c = 0; c < cutoff; ++c) {
        offsets.push_back(c * (base + 1));
        lens.push_back(base + 1);
    }
    if (base != 0)
        for (size_t c = 0; c + cutoff < threads; ++c) {
            offsets.push_back(split + c * base);
            lens.push_back(base);
        }
}

}  // namespace mh

using namespace mh;

// vanishing::Argument::sample_random (vanishing/prover.rs:41-75) after its
// seeds are drawn: a new device vector of 2^log_n values, chunk c of the
// parallelize split over `threads` workers drawn from
// ChaCha20Rng::from_seed(seeds + 32 c).  seed_count must equal the number of
// chunks (min(threads, 2^log_n)).
extern "C" RustError midnight_hip_fr_sample_random_poly(void **out, uint32_t log_n, size_t threads,
                                                        const uint8_t *seeds, size_t seed_count)
{
    if (!out || !seeds || threads == 0 || log_n > 28)
        return make_error((int)hipErrorInvalidValue, "invalid random polynomial arguments");
    *out = nullptr;
    const size_t n = size_t(1) << log_n;
    std::vector<size_t> off, len;
    parallelize_chunks(n, threads, off, len);
    if (seed_count != off.size())
        return make_error((int)hipErrorInvalidValue, "seed count differs from the number of chunks");
    MH_TRY
    hipStream_t st = cur_stream();
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = n;
    dv->values.alloc(n);
    random_fr_streams_device(dv->values.get(), off.data(), len.data(), seeds, off.size(), st);
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}
