// Field elements drawn from a ChaCha20Rng stream on the device: the
// vanishing argument's random polynomial (vanishing/prover.rs:41-75), n
// values Fq::random from ChaCha20Rng::from_seed(seed), bit-identical to
// drawing them one by one on the host.
//
// A fresh ChaCha20Rng only ever serves whole u64 pairs here, so candidate j
// (four next_u64) is words [8j, 8j + 8) of the keystream, i.e. one half of
// ChaCha block j / 2 (64-bit block counter, stream 0).  The top bit is
// cleared and the candidate is accepted when below r (about 90.6 % are);
// output i is the i-th accepted candidate.  Every lane makes four blocks
// (eight candidates); a count pass, a scan of the block counts and a
// placement pass (recomputing the blocks, which is cheaper than storing
// them) put the accepted values in order.
#include <vector>

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

__global__ __launch_bounds__(kBlock) void k_rng_count(ChaKey key, uint32_t *block_counts)
{
    __shared__ uint32_t sum;
    if (threadIdx.x == 0) sum = 0;
    __syncthreads();
    const size_t lane = blockIdx.x * (size_t)kBlock + threadIdx.x;
    atomicAdd(&sum, (uint32_t)__popc(lane_candidates(key, lane, nullptr)));
    __syncthreads();
    if (threadIdx.x == 0) block_counts[blockIdx.x] = sum;
}

// exclusive scan of the block counts, one lane (a few thousand blocks)
__global__ void k_rng_scan(uint32_t *counts, uint32_t nblocks)
{
    uint32_t run = 0;
    for (uint32_t b = 0; b < nblocks; ++b) {
        const uint32_t c = counts[b];
        counts[b] = run;
        run += c;
    }
    counts[nblocks] = run;
}

__global__ __launch_bounds__(kBlock) void k_rng_place(ChaKey key, const uint32_t *block_offsets, Fr *out, size_t n,
                                                      Fr r2)
{
    __shared__ uint32_t sh[2][kBlock];
    const size_t lane = blockIdx.x * (size_t)kBlock + threadIdx.x;
    Fr vals[kCandPerLane];
    const uint32_t flags = lane_candidates(key, lane, vals);
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
    size_t pos = block_offsets[blockIdx.x] + (t ? sh[b][t - 1] : 0u);
#pragma unroll
    for (int i = 0; i < kCandPerLane; ++i) {
        if (!((flags >> i) & 1)) continue;
        if (pos < n) store_fr(out + pos, vals[i] * r2);   // canonical -> Montgomery
        ++pos;
    }
}

}  // namespace

void random_fr_device(Fr *out, size_t n, const uint8_t seed[32], hipStream_t st)
{
    if (n == 0) return;
    ChaKey key;
    for (int i = 0; i < 8; ++i)
        key.k[i] = (uint32_t)seed[4 * i] | (uint32_t)seed[4 * i + 1] << 8 | (uint32_t)seed[4 * i + 2] << 16 |
                   (uint32_t)seed[4 * i + 3] << 24;
    // about 1.104 candidates per value; a wide margin (tens of standard
    // deviations at any n) and a check of the count
    const size_t want = n + n / 8 + 4096;
    const size_t nblocks = ceil_div(want, kCandPerBlock);
    MH_CHECK(nblocks < (1u << 31), (int)hipErrorInvalidValue, "too many random values");
    DevBuf<uint32_t> counts(nblocks + 1);
    hipLaunchKernelGGL(k_rng_count, dim3((unsigned)nblocks), dim3(kBlock), 0, st, key, counts.get());
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rng_scan, dim3(1), dim3(1), 0, st, counts.get(), (uint32_t)nblocks);
    MH_HIP(hipGetLastError());
    uint32_t total = 0;
    MH_HIP(hipMemcpyAsync(&total, counts.get() + nblocks, 4, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    MH_CHECK(total >= n, -1, "ChaCha20 candidates exhausted");
    Fr r2;
    std::memcpy(r2.v, HostFr::R2, 32);
    hipLaunchKernelGGL(k_rng_place, dim3((unsigned)nblocks), dim3(kBlock), 0, st, key, (const uint32_t *)counts.get(),
                       out, n, r2);
    MH_HIP(hipGetLastError());
}

}  // namespace mh
