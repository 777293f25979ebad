// Host-side primitives of the proof pipeline: Blake2b (RFC 7693) for the
// Fiat-Shamir transcript, the ChaCha20 RNG the prover draws blinding values
// from, field-element sampling and encodings, and the compressed G1
// encoding of commitments.
//
// Restated from the published algorithms the reference prover depends on:
//   - blake2b_simd 1.0.4: Blake2b-512, keyed / personalised parameter block
//     (RFC 7693 section 2.5);
//   - rand_core 0.6.4 / rand_chacha 0.3.1: seed_from_u64 expands the u64
//     with PCG32 (XSH-RR); ChaCha20 with 20 rounds, 64-bit block counter in
//     words 12-13, stream 0 in words 14-15, refilled four blocks at a time;
//     next_u64 reads two consecutive words (low first, across a refill when
//     one word is left), fill_bytes consumes whole words;
//   - Fq::random (vendor/midnight-curves-0.3.1/src/bls12_381/fq.rs:373-387):
//     four next_u64, top bit cleared, accepted when below r;
//   - Fq::from_uniform_bytes (fq.rs:704-726): 64 little-endian bytes
//     reduced mod r; to_repr: 32 little-endian canonical bytes;
//   - G1 to_bytes: 48-byte ZCash compressed encoding (big-endian x, flag
//     bits 0x80 compressed, 0x40 infinity, 0x20 y lexicographically largest).
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>

#include "host_fr.hpp"
#include "host_g1.hpp"

namespace mh {

// ---------------------------------------------------------------- Blake2b

class Blake2b {
public:
    Blake2b(size_t outlen, const uint8_t *key, size_t keylen, const uint8_t *personal16)
    {
        static constexpr uint64_t IV[8] = {0x6a09e667f3bcc908ULL, 0xbb67ae8584caa73bULL, 0x3c6ef372fe94f82bULL,
                                           0xa54ff53a5f1d36f1ULL, 0x510e527fade682d1ULL, 0x9b05688c2b3e6c1fULL,
                                           0x1f83d9abfb41bd6bULL, 0x5be0cd19137e2179ULL};
        for (int i = 0; i < 8; ++i) h_[i] = IV[i];
        outlen_ = outlen;
        h_[0] ^= 0x01010000ULL ^ ((uint64_t)keylen << 8) ^ (uint64_t)outlen;
        if (personal16) {
            h_[6] ^= le64(personal16);
            h_[7] ^= le64(personal16 + 8);
        }
        if (keylen) {
            uint8_t block[128] = {0};
            std::memcpy(block, key, keylen);
            update(block, 128);
        }
    }
    void update(const uint8_t *in, size_t n)
    {
        while (n > 0) {
            if (buflen_ == 128) {
                t_ += 128;
                compress(buf_, false);
                buflen_ = 0;
            }
            size_t take = 128 - buflen_ < n ? 128 - buflen_ : n;
            std::memcpy(buf_ + buflen_, in, take);
            buflen_ += take;
            in += take;
            n -= take;
        }
    }
    void update(const std::vector<uint8_t> &v) { update(v.data(), v.size()); }
    // Digest of everything absorbed so far; the state itself continues.
    std::vector<uint8_t> digest() const
    {
        Blake2b c = *this;
        c.t_ += c.buflen_;
        std::memset(c.buf_ + c.buflen_, 0, 128 - c.buflen_);
        c.compress(c.buf_, true);
        std::vector<uint8_t> out(outlen_);
        for (size_t i = 0; i < outlen_; ++i) out[i] = (uint8_t)(c.h_[i / 8] >> (8 * (i % 8)));
        return out;
    }

private:
    static uint64_t le64(const uint8_t *p)
    {
        uint64_t v = 0;
        for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
        return v;
    }
    static uint64_t rotr(uint64_t x, int r) { return (x >> r) | (x << (64 - r)); }
    void compress(const uint8_t *block, bool last)
    {
        static constexpr uint8_t S[12][16] = {
            {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
            {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
            {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
            {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
            {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0},
            {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3}};
        static constexpr uint64_t IV[8] = {0x6a09e667f3bcc908ULL, 0xbb67ae8584caa73bULL, 0x3c6ef372fe94f82bULL,
                                           0xa54ff53a5f1d36f1ULL, 0x510e527fade682d1ULL, 0x9b05688c2b3e6c1fULL,
                                           0x1f83d9abfb41bd6bULL, 0x5be0cd19137e2179ULL};
        uint64_t m[16], v[16];
        for (int i = 0; i < 16; ++i) m[i] = le64(block + 8 * i);
        for (int i = 0; i < 8; ++i) {
            v[i] = h_[i];
            v[i + 8] = IV[i];
        }
        v[12] ^= t_;
        if (last) v[14] = ~v[14];
        auto G = [&](int a, int b, int c, int d, uint64_t x, uint64_t y) {
            v[a] = v[a] + v[b] + x;
            v[d] = rotr(v[d] ^ v[a], 32);
            v[c] = v[c] + v[d];
            v[b] = rotr(v[b] ^ v[c], 24);
            v[a] = v[a] + v[b] + y;
            v[d] = rotr(v[d] ^ v[a], 16);
            v[c] = v[c] + v[d];
            v[b] = rotr(v[b] ^ v[c], 63);
        };
        for (int r = 0; r < 12; ++r) {
            const uint8_t *s = S[r];
            G(0, 4, 8, 12, m[s[0]], m[s[1]]);
            G(1, 5, 9, 13, m[s[2]], m[s[3]]);
            G(2, 6, 10, 14, m[s[4]], m[s[5]]);
            G(3, 7, 11, 15, m[s[6]], m[s[7]]);
            G(0, 5, 10, 15, m[s[8]], m[s[9]]);
            G(1, 6, 11, 12, m[s[10]], m[s[11]]);
            G(2, 7, 8, 13, m[s[12]], m[s[13]]);
            G(3, 4, 9, 14, m[s[14]], m[s[15]]);
        }
        for (int i = 0; i < 8; ++i) h_[i] ^= v[i] ^ v[i + 8];
    }

    uint64_t h_[8];
    uint64_t t_ = 0;   // bytes compressed (messages here stay below 2^64 bytes)
    uint8_t buf_[128];
    size_t buflen_ = 0;
    size_t outlen_ = 64;
};

// ---------------------------------------------------------------- ChaCha20 RNG

class ChaChaRng {
public:
    static ChaChaRng from_seed(const uint8_t seed[32])
    {
        ChaChaRng r;
        for (int i = 0; i < 8; ++i)
            r.key_[i] = (uint32_t)seed[4 * i] | (uint32_t)seed[4 * i + 1] << 8 | (uint32_t)seed[4 * i + 2] << 16 |
                        (uint32_t)seed[4 * i + 3] << 24;
        return r;
    }
    static ChaChaRng seed_from_u64(uint64_t state)
    {
        const uint64_t MUL = 6364136223846793005ULL, INC = 11634580027462260723ULL;
        uint8_t seed[32];
        for (int i = 0; i < 8; ++i) {
            state = state * MUL + INC;
            const uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
            const uint32_t rot = (uint32_t)(state >> 59);
            const uint32_t x = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
            for (int b = 0; b < 4; ++b) seed[4 * i + b] = (uint8_t)(x >> (8 * b));
        }
        return from_seed(seed);
    }
    uint32_t next_u32()
    {
        if (idx_ >= 64) refill();
        return buf_[idx_++];
    }
    uint64_t next_u64()
    {
        if (idx_ < 63) {
            const uint64_t v = (uint64_t)buf_[idx_] | (uint64_t)buf_[idx_ + 1] << 32;
            idx_ += 2;
            return v;
        }
        if (idx_ >= 64) {
            refill();
            idx_ = 2;
            return (uint64_t)buf_[0] | (uint64_t)buf_[1] << 32;
        }
        const uint64_t lo = buf_[63];
        refill();
        idx_ = 1;
        return lo | (uint64_t)buf_[0] << 32;
    }
    void fill_bytes(uint8_t *dst, size_t n)
    {
        size_t done = 0;
        while (done < n) {
            if (idx_ >= 64) refill();
            const uint32_t w = buf_[idx_++];
            for (int b = 0; b < 4 && done < n; ++b) dst[done++] = (uint8_t)(w >> (8 * b));
        }
    }
    // Fq::random: rejection sampling of 255-bit values
    HostFr random_fr()
    {
        for (;;) {
            uint64_t raw[4];
            for (int i = 0; i < 4; ++i) raw[i] = next_u64();
            raw[3] &= 0x7fffffffffffffffULL;
            if (!HostFr::geq_p(raw)) return HostFr::from_canonical(raw);
        }
    }

private:
    static uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
    void block(uint32_t *out, uint64_t counter) const
    {
        uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        std::memcpy(s + 4, key_, 32);
        s[12] = (uint32_t)counter;
        s[13] = (uint32_t)(counter >> 32);
        s[14] = 0;
        s[15] = 0;
        uint32_t x[16];
        std::memcpy(x, s, sizeof(x));
        auto qr = [&](int a, int b, int c, int d) {
            x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 16);
            x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 12);
            x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 8);
            x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 7);
        };
        for (int i = 0; i < 10; ++i) {
            qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
            qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
        }
        for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
    }
    void refill()
    {
        for (int b = 0; b < 4; ++b) block(buf_ + 16 * b, counter_ + b);
        counter_ += 4;
        idx_ = 0;
    }
    uint32_t key_[8] = {0};
    uint64_t counter_ = 0;
    uint32_t buf_[64] = {0};
    uint32_t idx_ = 64;
};

// ---------------------------------------------------------------- encodings

// Fq::from_uniform_bytes: (lo + hi * 2^256) mod r, in Montgomery form
inline HostFr fr_from_uniform_bytes(const uint8_t *b)
{
    uint64_t lo[4], hi[4];
    for (int i = 0; i < 4; ++i) {
        lo[i] = 0;
        hi[i] = 0;
        for (int j = 7; j >= 0; --j) {
            lo[i] = (lo[i] << 8) | b[8 * i + j];
            hi[i] = (hi[i] << 8) | b[32 + 8 * i + j];
        }
    }
    // Montgomery product mont(a, b) = a b / R.  The Montgomery form of lo is
    // mont(lo, R^2); that of hi * 2^256 = hi * R is mont(hi, R^3), and the
    // raw integer R^3 mod r is mont(R^2, R^2).
    const HostFr l{{lo[0], lo[1], lo[2], lo[3]}}, h{{hi[0], hi[1], hi[2], hi[3]}};
    const HostFr r2{{HostFr::R2[0], HostFr::R2[1], HostFr::R2[2], HostFr::R2[3]}};
    const HostFr r3 = r2 * r2;
    return l * r2 + h * r3;
}

// Fq::to_repr: 32 canonical little-endian bytes
inline void fr_to_bytes(const HostFr &a, uint8_t out[32])
{
    uint64_t c[4];
    a.to_canonical(c);
    for (int i = 0; i < 32; ++i) out[i] = (uint8_t)(c[i / 8] >> (8 * (i % 8)));
}

// 48-byte compressed encoding of a Jacobian point (Montgomery limbs)
void g1_compress(const RawG1Projective &p, uint8_t out[48]);

}  // namespace mh
