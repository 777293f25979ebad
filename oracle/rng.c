/*
 * Oracle: the seeded RNG that synthetic workloads are drawn from.
 * TEST INFRASTRUCTURE ONLY.
 *
 * The reference draws test inputs with ChaCha20Rng::seed_from_u64(seed)
 * (e.g. metal-fixed-base-msm-regression/tests/kzg_commit_batch.rs:40-44)
 * and field elements with Fq::random (midnight-curves fq.rs:373-387).
 * Those crates (rand_core 0.6.4, rand_chacha 0.3.1) are not vendored in the
 * reference; their published algorithms are restated here:
 *   - seed_from_u64: PCG32 (XSH-RR) expands the u64 into the 32-byte seed,
 *     advancing the state before each output;
 *   - ChaCha20: DJB ChaCha with 20 rounds, 64-bit block counter in words
 *     12-13, stream id 0 in words 14-15, refilled four blocks at a time;
 *   - next_u64: two consecutive u32 words, low word first.
 * Pinned by:
 *   - the RFC 7539 section 2.3.2 block-function test vector (block function,
 *     key/counter/nonce word layout);
 *   - the RFC 7539 appendix A.1 keystream vectors #1 and #2 (all-zero key and
 *     nonce, block counters 0 and 1) read through the stream: from_seed of a
 *     zero seed yields exactly those 32 words, next_u64 pairs them low word
 *     first, and a u64 straddling the 64-word refill takes word 63 as its low
 *     half and the next refill's word 0 as its high half;
 *   - O'Neill's pcg32 demo outputs (pcg32_srandom_r(42, 54)) for the LCG step
 *     and the XSH-RR output permutation seed_from_u64 is built from.
 * Still restated only from rand_core 0.6.4's published source, with no
 * vector anywhere to check them against: seed_from_u64's increment
 * 11634580027462260723 and its advance-then-output order (parity unpinned
 * for seed_from_u64's exact seed bytes; every draw from a given 32-byte seed
 * is pinned by the vectors above).
 */
#include "oracle.h"
#include <string.h>

static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

#define QR(a, b, c, d)                 \
    a += b; d ^= a; d = rotl32(d, 16); \
    c += d; b ^= c; b = rotl32(b, 12); \
    a += b; d ^= a; d = rotl32(d, 8);  \
    c += d; b ^= c; b = rotl32(b, 7);

void orc_chacha20_block(uint32_t out[16], const uint32_t key[8], uint64_t counter,
                        const uint32_t nonce[2])
{
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    memcpy(s + 4, key, 32);
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = nonce[0];
    s[15] = nonce[1];
    uint32_t x[16];
    memcpy(x, s, sizeof(x));
    for (int i = 0; i < 10; ++i) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

static void refill(orc_rng *r)
{
    static const uint32_t zero_nonce[2] = {0, 0};
    for (int b = 0; b < 4; ++b) orc_chacha20_block(r->buf + 16 * b, r->key, r->counter + b, zero_nonce);
    r->counter += 4;
    r->idx = 0;
}

void orc_rng_from_seed(orc_rng *r, const uint8_t seed[32])
{
    for (int i = 0; i < 8; ++i)
        r->key[i] = (uint32_t)seed[4 * i] | (uint32_t)seed[4 * i + 1] << 8 |
                    (uint32_t)seed[4 * i + 2] << 16 | (uint32_t)seed[4 * i + 3] << 24;
    r->counter = 0;
    r->idx = 64;
}

/* PCG32's 64-bit LCG step with increment inc (the multiplier of pcg32) */
uint64_t orc_pcg32_advance(uint64_t state, uint64_t inc)
{
    return state * 6364136223846793005ULL + inc;
}

/* PCG32's XSH-RR output permutation of a 64-bit state */
uint32_t orc_pcg32_output(uint64_t state)
{
    uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
    uint32_t rot = (uint32_t)(state >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}

void orc_rng_seed_from_u64(orc_rng *r, uint64_t state)
{
    const uint64_t INC = 11634580027462260723ULL;
    uint8_t seed[32];
    for (int i = 0; i < 8; ++i) {
        /* rand_core advances the state before each output */
        state = orc_pcg32_advance(state, INC);
        uint32_t x = orc_pcg32_output(state);
        seed[4 * i] = (uint8_t)x;
        seed[4 * i + 1] = (uint8_t)(x >> 8);
        seed[4 * i + 2] = (uint8_t)(x >> 16);
        seed[4 * i + 3] = (uint8_t)(x >> 24);
    }
    orc_rng_from_seed(r, seed);
}

uint32_t orc_rng_next_u32(orc_rng *r)
{
    if (r->idx >= 64) refill(r);
    return r->buf[r->idx++];
}

uint64_t orc_rng_next_u64(orc_rng *r)
{
    if (r->idx < 63) {
        uint64_t v = (uint64_t)r->buf[r->idx] | (uint64_t)r->buf[r->idx + 1] << 32;
        r->idx += 2;
        return v;
    }
    if (r->idx >= 64) {
        refill(r);
        r->idx = 2;
        return (uint64_t)r->buf[0] | (uint64_t)r->buf[1] << 32;
    }
    /* one word left: low half from it, high half from the next refill */
    uint64_t lo = r->buf[63];
    refill(r);
    r->idx = 1;
    return lo | (uint64_t)r->buf[0] << 32;
}

void orc_fr_random(orc_fr *out, orc_rng *rng)
{
    for (;;) {
        uint64_t raw[4];
        for (int i = 0; i < 4; ++i) raw[i] = orc_rng_next_u64(rng);
        raw[3] &= 0xffffffffffffffffULL >> 1;   /* 256 - NUM_BITS(255) */
        if (orc_fr_from_canonical(out, raw)) return;
    }
}

void orc_fr_random_many(orc_fr *out, size_t n, orc_rng *rng)
{
    for (size_t i = 0; i < n; ++i) orc_fr_random(&out[i], rng);
}
