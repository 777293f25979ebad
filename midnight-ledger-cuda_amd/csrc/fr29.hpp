// BLS12-381 Fr products in radix 2^29 for the NTT.
//
// Nine 29-bit limbs (261 bits).  The product is CIOS Montgomery
// multiplication with 64-bit lazy accumulators, as the radix-2^28 Fp code
// does it: a 29 x 29-bit product is below 2^58 and every accumulator
// collects at most 18 of them (< 2^62.2), so each product is one
// v_mad_u64_u32 with the accumulator as addend -- no carry flags, no zero
// high words.  r = 1 mod 2^32 makes the per-row quotient m = -t0 mod 2^29
// (no multiply) and the lowest limb of r equal to 1 (no multiply).
//
// The Montgomery radix of this product is 2^261, not the 2^256 of the rest
// of the code.  The NTT multiplies a data value in the usual form a 2^256
// by a twiddle stored as w 2^261 mod r:
//   mont29(a 2^256, w 2^261) = a w 2^256 2^261 / 2^261 = (a w) 2^256,
// exactly the usual form of the product, so data never changes
// representation -- only the twiddle tables (and the NTT's constant
// factors) are scaled by 2^5 once on the host.  The result is fully
// reduced, so it is limb-identical to the 32-bit product.
#pragma once
#include "fields.hpp"

namespace mh {

struct Fr29 {
    uint32_t v[9];
};

constexpr uint32_t kM29 = 0x1fffffffu;

__device__ __forceinline__ uint32_t r29(int i)
{
    constexpr uint32_t R[9] = {0x00000001u, 0x1ffffff8u, 0x1f96ffbfu, 0x1b4805ffu, 0x1d80553bu,
                               0x0c0404d0u, 0x1520cce7u, 0x0a6533afu, 0x0073eda7u};
    return R[i];
}

// same integer, 8 x 32-bit limbs -> 9 x 29-bit limbs
__device__ __forceinline__ Fr29 fr29_from32(const Fr &x)
{
    Fr29 r;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int bit = 29 * j, q = bit >> 5, s = bit & 31;
        const uint32_t lo = x.v[q];
        const uint32_t hi = q + 1 < 8 ? x.v[q + 1] : 0u;
        r.v[j] = (uint32_t)((((uint64_t)hi << 32) | lo) >> s) & kM29;
    }
    return r;
}

// 9 x 29-bit limbs (value below 2^256) -> 8 x 32-bit limbs
__device__ __forceinline__ Fr fr29_to32(const Fr29 &x)
{
    Fr r;
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        const int bit = 32 * w, j = bit / 29, s = bit % 29;
        uint64_t acc = (uint64_t)(x.v[j] >> s);
        int have = 29 - s;
        if (j + 1 < 9) acc |= (uint64_t)x.v[j + 1] << have;
        if (have + 29 < 32 && j + 2 < 9) acc |= (uint64_t)x.v[j + 2] << (have + 29);
        r.v[w] = (uint32_t)acc;
    }
    return r;
}

// r - modulus if r >= modulus (r < 2 modulus, limbs below 2^29)
__device__ __forceinline__ void fr29_reduce(Fr29 &r)
{
    uint32_t d[9];
    int32_t bw = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        int32_t x = (int32_t)r.v[j] - (int32_t)r29(j) - bw;
        bw = (int32_t)((uint32_t)x >> 31);
        d[j] = (uint32_t)x & kM29;
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) r.v[j] = bw ? r.v[j] : d[j];
}

// a + b mod r for fully reduced a, b (any Montgomery radix: the sum of two
// representations is the representation of the sum)
__device__ __forceinline__ Fr29 fr29_add(const Fr29 &a, const Fr29 &b)
{
    Fr29 r;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const uint32_t s = a.v[j] + b.v[j] + c;
        c = s >> 29;
        r.v[j] = s & kM29;
    }
    fr29_reduce(r);
    return r;
}

// a * b / 2^261 mod r, fully reduced; a < 2^261 (any 9-limb value), b < r
// (then the running value stays below a b / 2^261 + r < 2r).
__device__ __forceinline__ Fr29 fr29_mul(const Fr29 &a, const Fr29 &b)
{
    uint64_t t[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
#pragma unroll
        for (int j = 0; j < 9; ++j) t[j] += (uint64_t)a.v[j] * b.v[i];
        const uint32_t m = (0u - (uint32_t)t[0]) & kM29;
        t[0] += m;   // r_0 = 1
#pragma unroll
        for (int j = 1; j < 9; ++j) t[j] += (uint64_t)m * r29(j);
        // t[0] is a multiple of 2^29: shift the row down one limb
        const uint64_t carry = t[0] >> 29;
#pragma unroll
        for (int j = 0; j < 8; ++j) t[j] = t[j + 1];
        t[0] += carry;
        t[8] = 0;
    }
    Fr29 r;
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        c += t[j];
        r.v[j] = (uint32_t)c & kM29;
        c >>= 29;
    }
    fr29_reduce(r);   // r < 2 modulus
    return r;
}

// a - b mod r for fully reduced a, b
__device__ __forceinline__ Fr29 fr29_sub(const Fr29 &a, const Fr29 &b)
{
    Fr29 r;
    int32_t bw = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int32_t x = (int32_t)a.v[j] - (int32_t)b.v[j] - bw;
        bw = (int32_t)((uint32_t)x >> 31);
        r.v[j] = (uint32_t)x & kM29;
    }
    // a borrow out of the top limb: the difference wrapped by 2^261; add r
    const uint32_t mask = 0u - (uint32_t)bw;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const uint32_t s = r.v[j] + (r29(j) & mask) + c;
        c = s >> 29;
        r.v[j] = s & kM29;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Loose values: any representative below 2r (limbs below 2^29).  Products of
// loose values skip the final subtraction: for a, b < 2r the running value is
// below a b / 2^261 + r < 4r^2 / 2^261 + r < 1.06 r.  Sums and differences
// are brought back below 2r by one conditional subtraction of 2r.  A loose
// value is fully reduced by one fr29_reduce.

__device__ __forceinline__ uint32_t r2x29(int i)   // 2r in 29-bit limbs
{
    constexpr uint32_t R2[9] = {0x00000002u, 0x1ffffff0u, 0x1f2dff7fu, 0x16900bffu, 0x1b00aa77u,
                                0x180809a1u, 0x0a4199ceu, 0x14ca675fu, 0x00e7db4eu};
    return R2[i];
}

// r - 2r if r >= 2r (r < 4r, limbs below 2^29)
__device__ __forceinline__ void fr29_reduce_2r(Fr29 &r)
{
    uint32_t d[9];
    int32_t bw = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        int32_t x = (int32_t)r.v[j] - (int32_t)r2x29(j) - bw;
        bw = (int32_t)((uint32_t)x >> 31);
        d[j] = (uint32_t)x & kM29;
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) r.v[j] = bw ? r.v[j] : d[j];
}

// a * b / 2^261 mod r for loose a, b; loose result
__device__ __forceinline__ Fr29 fr29_mul_loose(const Fr29 &a, const Fr29 &b)
{
    uint64_t t[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
#pragma unroll
        for (int j = 0; j < 9; ++j) t[j] += (uint64_t)a.v[j] * b.v[i];
        const uint32_t m = (0u - (uint32_t)t[0]) & kM29;
        t[0] += m;   // r_0 = 1
#pragma unroll
        for (int j = 1; j < 9; ++j) t[j] += (uint64_t)m * r29(j);
        const uint64_t carry = t[0] >> 29;
#pragma unroll
        for (int j = 0; j < 8; ++j) t[j] = t[j + 1];
        t[0] += carry;
        t[8] = 0;
    }
    Fr29 r;
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        c += t[j];
        r.v[j] = (uint32_t)c & kM29;
        c >>= 29;
    }
    return r;
}

// (a b + c d) / 2^261 mod r for loose a, b, c, d with one reduction: each
// row adds a_j b_i + c_j d_i, then m r; loose result (below (8r^2) / 2^261 +
// r < 1.12 r), accumulators below 27 products of 2^58.
__device__ __forceinline__ Fr29 fr29_mul2_loose(const Fr29 &a, const Fr29 &b, const Fr29 &c, const Fr29 &d)
{
    uint64_t t[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
#pragma unroll
        for (int j = 0; j < 9; ++j) t[j] += (uint64_t)a.v[j] * b.v[i];
#pragma unroll
        for (int j = 0; j < 9; ++j) t[j] += (uint64_t)c.v[j] * d.v[i];
        const uint32_t m = (0u - (uint32_t)t[0]) & kM29;
        t[0] += m;   // r_0 = 1
#pragma unroll
        for (int j = 1; j < 9; ++j) t[j] += (uint64_t)m * r29(j);
        const uint64_t carry = t[0] >> 29;
#pragma unroll
        for (int j = 0; j < 8; ++j) t[j] = t[j + 1];
        t[0] += carry;
        t[8] = 0;
    }
    Fr29 r;
    uint64_t cc = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        cc += t[j];
        r.v[j] = (uint32_t)cc & kM29;
        cc >>= 29;
    }
    return r;
}

// a + b for loose a, b; loose result
__device__ __forceinline__ Fr29 fr29_add_loose(const Fr29 &a, const Fr29 &b)
{
    Fr29 r;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const uint32_t s = a.v[j] + b.v[j] + c;
        c = s >> 29;
        r.v[j] = s & kM29;
    }
    fr29_reduce_2r(r);
    return r;
}

// a - b (as a - b + 2r) for loose a, b; loose result
__device__ __forceinline__ Fr29 fr29_sub_loose(const Fr29 &a, const Fr29 &b)
{
    Fr29 r;
    int32_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int32_t s = (int32_t)a.v[j] + (int32_t)r2x29(j) - (int32_t)b.v[j] + c;
        c = s >> 29;   // arithmetic: -1, 0 or 1
        r.v[j] = (uint32_t)s & kM29;
    }
    fr29_reduce_2r(r);
    return r;
}

// ---------------------------------------------------------------------------
// Lazy values: limbs may exceed 29 bits (each below 2^31 where a value enters
// a product, below 2^32 otherwise) and the value may exceed 2r; only the
// integer's bound matters, tracked by the caller.  fr29_mul_loose accepts
// such a value as either operand when its bound times the other's stays
// below r 2^261 (every accumulator then stays below 2^64), and returns a
// normalized value below a b / 2^261 + r.

// a + b, no carries
__device__ __forceinline__ Fr29 fr29_lazy_add(const Fr29 &a, const Fr29 &b)
{
    Fr29 r;
#pragma unroll
    for (int j = 0; j < 9; ++j) r.v[j] = a.v[j] + b.v[j];
    return r;
}

// 4r with 2^29 lent to every limb below the top (each of those limbs is at
// least 2^29, the top one at least the top limb of any value below 2r), so
// that a - b + 4r can be formed limb by limb for a normalized b < 2r
__device__ __forceinline__ uint32_t r4x29_lent(int i)
{
    constexpr uint32_t O[9] = {0x20000004u, 0x3fffffdfu, 0x3e5bfefeu, 0x2d2017feu, 0x360154eeu,
                               0x30101342u, 0x3483339cu, 0x2994cebdu, 0x01cfb69cu};
    return O[i];
}

// a - b + 4r for b normalized and below 2r, no carries
__device__ __forceinline__ Fr29 fr29_lazy_sub4r(const Fr29 &a, const Fr29 &b)
{
    Fr29 r;
#pragma unroll
    for (int j = 0; j < 9; ++j) r.v[j] = a.v[j] + r4x29_lent(j) - b.v[j];
    return r;
}

// 8r with 2^30 lent to every limb below the top (each of those limbs is at
// least 2^30 - 2, the top one at least twice the top limb of any value
// below 2r), so that a - b + 8r can be formed limb by limb for b the lazy
// sum of two normalized values below 2r
__device__ __forceinline__ uint32_t r8x29_lent(int i)
{
    constexpr uint32_t O[9] = {0x40000008u, 0x5fffffbeu, 0x5cb7fdfdu, 0x5a402ffdu, 0x4c02a9dcu,
                               0x40202685u, 0x49066739u, 0x53299d7bu, 0x039f6d38u};
    return O[i];
}

// a - b + 8r for b = x + y, x and y normalized and below 2r (limbs below
// 2^30, value below 4r), no carries; a's limbs below 2^30 keep the result's
// below 2^32
__device__ __forceinline__ Fr29 fr29_lazy_sub8r(const Fr29 &a, const Fr29 &b)
{
    Fr29 r;
#pragma unroll
    for (int j = 0; j < 9; ++j) r.v[j] = a.v[j] + r8x29_lent(j) - b.v[j];
    return r;
}

// carry propagation: limbs below 2^32, value below 2^261 -> limbs below 2^29
__device__ __forceinline__ void fr29_norm(Fr29 &a)
{
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t s = a.v[j] + c;
        c = s >> 29;
        a.v[j] = s & kM29;
    }
    a.v[8] += c;
}

// normalized x below 2^261 -> the same class below 2r: x - q r with
// q = floor(x8 M / 2^48), M = floor(2^48 / (r8 + 1)), r8 the top limb of r.
// x8 M / 2^48 is below x / r and short of it by at most (x / r) / r8 plus
// 2^-23 (the two floors), under 10^-5 for x / r < 70.6, so
// floor(x / r) - 1 <= q <= floor(x / r)
__device__ __forceinline__ void fr29_reduce_top(Fr29 &a)
{
    const uint32_t q = (uint32_t)(((uint64_t)a.v[8] * 0x2355091ull) >> 48);
    int64_t carry = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int64_t t = (int64_t)a.v[j] - (int64_t)q * (int64_t)r29(j) + carry;
        a.v[j] = (uint32_t)t & kM29;
        carry = t >> 29;
    }
}

// ---------------------------------------------------------------------------
// The scaled form: a value v kept as v 2^261 mod r (fully reduced, 9 limbs),
// the form in which fr29_mul is closed: mont29(a 2^261, b 2^261) = (a b)
// 2^261.  Sums and differences of scaled values are scaled values.  Data in
// the usual form (v 2^256, 8 x 32 bits) enters by a multiplication by 32
// (a 5-bit shift and a reduction) and leaves by one product with 2^256 mod r.

// 2^261 mod r: the scaled form of one
__device__ __forceinline__ Fr29 s29_one()
{
    constexpr uint32_t O[9] = {0x1fffffbau, 0x0000022fu, 0x1cb61180u, 0x0a4e5c00u, 0x0ee8b1a2u,
                               0x16e6aedfu, 0x1907f8bbu, 0x0853ddf7u, 0x004d043fu};
    Fr29 r;
#pragma unroll
    for (int j = 0; j < 9; ++j) r.v[j] = O[j];
    return r;
}

// usual form (8 x 32, fully reduced) -> scaled form: 32 x mod r
__device__ __forceinline__ Fr29 s29_from_usual(const Fr &x)
{
    // V = 32 x < 32 r < 2^260: bits [29 j - 5, 29 j + 24) of x in limb j
    Fr29 v;
    v.v[0] = (x.v[0] << 5) & kM29;
#pragma unroll
    for (int j = 1; j < 9; ++j) {
        const int bit = 29 * j - 5, q = bit >> 5, s = bit & 31;
        const uint32_t lo = x.v[q];
        const uint32_t hi = q + 1 < 8 ? x.v[q + 1] : 0u;
        v.v[j] = (uint32_t)((((uint64_t)hi << 32) | lo) >> s) & kM29;
    }
    // q = floor(v8 M / 2^48) with M = floor(2^48 / (r8 + 1)), r8 the top
    // limb of r: floor(V / r) - 2 <= q <= floor(V / r), so V - q r < 3 r
    const uint32_t q = (uint32_t)(((uint64_t)v.v[8] * 0x2355091ull) >> 48);
    int64_t carry = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int64_t t = (int64_t)v.v[j] - (int64_t)q * (int64_t)r29(j) + carry;
        v.v[j] = (uint32_t)t & kM29;
        carry = t >> 29;   // arithmetic: the borrow into the next limb
    }
    fr29_reduce(v);
    fr29_reduce(v);
    return v;
}

// scaled form (loose or fully reduced) -> usual form (8 x 32, fully reduced):
// one product by 2^256 mod r
__device__ __forceinline__ Fr s29_to_usual(const Fr29 &v)
{
    return fr29_to32(fr29_mul(v, fr29_from32(Fr::one())));
}

// (a w) in the usual form, for a in the usual form (8 x 32) and w stored as
// w 2^261 mod r (8 x 32, fully reduced)
__device__ __forceinline__ Fr fr_mul_scaled(const Fr &a, const Fr &w261)
{
    return fr29_to32(fr29_mul(fr29_from32(a), fr29_from32(w261)));
}

// the same with w already repacked (one twiddle, several products)
__device__ __forceinline__ Fr fr_mul_scaled29(const Fr &a, const Fr29 &w261)
{
    return fr29_to32(fr29_mul(fr29_from32(a), w261));
}

}  // namespace mh
