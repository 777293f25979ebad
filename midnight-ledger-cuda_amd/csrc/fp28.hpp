// BLS12-381 Fp in radix 2^28 for the MSM accumulation loop.
//
// Fourteen 28-bit limbs (392 bits) in Montgomery form with R = 2^392.  The
// product is finely integrated product scanning (FIPS; Koc, Acar, Kaliski
// 1996): every column adds 28-bit x 28-bit products into one 64-bit
// accumulator, which v_mad_u64_u32 does in a single instruction per
// product with the accumulator itself as the 64-bit addend -- no carry flag,
// no zero-extension.  A column holds at most 28 products below 2^56, so the
// accumulator cannot overflow.  With 32-bit limbs every product also needs a
// carry-propagating add (and a zero high word), about 1.6x the instructions.
// Values are kept fully reduced in [0, p), so zero tests and equality are
// limb comparisons.
//
// Conversions to and from the 32-bit, R = 2^384 form used everywhere else
// (the reference's RawG1Affine layout) are a Montgomery product by
// 2^400 mod p (in) or 2^384 mod p (out) plus a bit repacking.
#pragma once
#include "curve.hpp"
#include "fields.hpp"

namespace mh {

struct Fp28 {
    uint32_t v[14];
};

constexpr uint32_t kM28 = 0x0fffffffu;
constexpr uint32_t kPinv28 = 0x0ffcfffdu;   // -p^-1 mod 2^28

__device__ __forceinline__ uint32_t p28(int i)
{
    constexpr uint32_t P[14] = {0xfffaaabu, 0xfefffffu, 0x3ffffb9u, 0xfffeb15u, 0x6241eabu, 0xa0f6b0fu, 0xf6730d2u,
                                0xf38512bu, 0x4774b84u, 0x4bacd76u, 0xba7b643u, 0xe69a4b1u, 0x1ea397fu, 0x001a011u};
    return P[i];
}

__device__ __forceinline__ Fp28 fp28_const(int which)
{
    // 0: 2^400 mod p (std -> internal), 1: 2^384 mod p (internal -> std)
    constexpr uint32_t TO_INT[14] = {0x80e6299u, 0x3500034u, 0xeb12856u, 0xdeb2699u, 0xc988670u,
                                     0x4ef6697u, 0x70983e8u, 0xa4e6fe9u, 0x3e8a053u, 0xecf271eu,
                                     0xc20d323u, 0x6eb6385u, 0x47f1286u, 0x00156dau};
    constexpr uint32_t TO_STD[14] = {0x002fffdu, 0x0900000u, 0xc000276u, 0x000bc40u, 0x8baebf4u,
                                     0x5753c75u, 0x55f4898u, 0x7052574u, 0x7ce5853u, 0x56ec6d7u,
                                     0x71a97a2u, 0xe4935c0u, 0xec3fa80u, 0x0015f65u};
    Fp28 r;
#pragma unroll
    for (int i = 0; i < 14; ++i) r.v[i] = which == 0 ? TO_INT[i] : TO_STD[i];
    return r;
}

__device__ __forceinline__ Fp28 fp28_one()
{
    constexpr uint32_t O[14] = {0x347fcb8u, 0xd800000u, 0x002b119u, 0x0cde6d2u, 0xc7212e0u, 0x83a2090u, 0x037669fu,
                                0xda0f73eu, 0x9b09b42u, 0x1297bb0u, 0x515d98fu, 0x012ca7cu, 0x659fcfau, 0x000577au};
    Fp28 r;
#pragma unroll
    for (int i = 0; i < 14; ++i) r.v[i] = O[i];
    return r;
}

__device__ __forceinline__ Fp28 fp28_zero()
{
    Fp28 r;
#pragma unroll
    for (int i = 0; i < 14; ++i) r.v[i] = 0;
    return r;
}

__device__ __forceinline__ bool fp28_is_zero(const Fp28 &a)
{
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) o |= a.v[i];
    return o == 0;
}

// x - p if x >= p (x < 2p, limbs < 2^28 except the top, which is < 2^19)
__device__ __forceinline__ void fp28_reduce(Fp28 &x)
{
    uint32_t d[14];
    int32_t bw = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        int32_t t = (int32_t)x.v[i] - (int32_t)p28(i) - bw;
        bw = (int32_t)((uint32_t)t >> 31);
        d[i] = (uint32_t)t & kM28;
    }
#pragma unroll
    for (int i = 0; i < 14; ++i) x.v[i] = bw ? x.v[i] : d[i];
}

// Row-oriented form (CIOS order) with 64-bit lazy accumulators: row i adds
// a_j * b_i and m * p_j into fourteen independent accumulators (28
// independent multiply-adds), so consecutive instructions do not wait on
// one another; only the 28-bit carry of the lowest word moves up per row.
// Each accumulator collects at most 28 products below 2^56 (< 2^61).
template <bool kReduce = true>
__device__ __forceinline__ Fp28 fp28_mul(const Fp28 &a, const Fp28 &b)
{
    uint64_t t[14];
#pragma unroll
    for (int j = 0; j < 14; ++j) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
#pragma unroll
        for (int j = 0; j < 14; ++j) t[j] += (uint64_t)a.v[j] * b.v[i];
        const uint32_t m = ((uint32_t)t[0] * kPinv28) & kM28;
#pragma unroll
        for (int j = 0; j < 14; ++j) t[j] += (uint64_t)m * p28(j);
        // t[0] is now a multiple of 2^28: shift the row down one limb
        const uint64_t carry = t[0] >> 28;
#pragma unroll
        for (int j = 0; j < 13; ++j) t[j] = t[j + 1];
        t[0] += carry;
        t[13] = 0;
    }
    Fp28 r;
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 14; ++j) {
        c += t[j];
        r.v[j] = (uint32_t)c & kM28;
        c >>= 28;
    }
    if (kReduce) fp28_reduce(r);
    return r;
}

// (a b + c d) / R mod p with one Montgomery reduction for both products:
// each row adds a_j b_i + c_j d_i, then m p.  For operands below 2p the
// result is below (4p^2 + 4p^2) / 2^392 + p < 2p (loose), and an accumulator
// collects at most 42 products below 2^56 (< 2^62).  Saves the second
// product's reduction (14 of its 28 multiply-adds per row).
__device__ __forceinline__ Fp28 fp28_mul2_lz(const Fp28 &a, const Fp28 &b, const Fp28 &c, const Fp28 &d)
{
    uint64_t t[14];
#pragma unroll
    for (int j = 0; j < 14; ++j) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
#pragma unroll
        for (int j = 0; j < 14; ++j) t[j] += (uint64_t)a.v[j] * b.v[i];
#pragma unroll
        for (int j = 0; j < 14; ++j) t[j] += (uint64_t)c.v[j] * d.v[i];
        const uint32_t m = ((uint32_t)t[0] * kPinv28) & kM28;
#pragma unroll
        for (int j = 0; j < 14; ++j) t[j] += (uint64_t)m * p28(j);
        const uint64_t carry = t[0] >> 28;
#pragma unroll
        for (int j = 0; j < 13; ++j) t[j] = t[j + 1];
        t[0] += carry;
        t[13] = 0;
    }
    Fp28 r;
    uint64_t cc = 0;
#pragma unroll
    for (int j = 0; j < 14; ++j) {
        cc += t[j];
        r.v[j] = (uint32_t)cc & kM28;
        cc >>= 28;
    }
    return r;
}

// Column-oriented form (finely integrated product scanning), kept for
// squaring where the symmetric cross products halve the multiplies.
__device__ __forceinline__ Fp28 fp28_mul_fips(const Fp28 &a, const Fp28 &b)
{
    uint64_t acc = 0;
    uint32_t m[14];
#pragma unroll
    for (int k = 0; k < 14; ++k) {
        uint64_t e = 0, o = 0, q = 0;
#pragma unroll
        for (int i = 0; i <= k; ++i) {
            if (i & 1) o += (uint64_t)a.v[i] * b.v[k - i];
            else e += (uint64_t)a.v[i] * b.v[k - i];
        }
#pragma unroll
        for (int i = 0; i < k; ++i) q += (uint64_t)m[i] * p28(k - i);
        acc += e + o + q;
        m[k] = ((uint32_t)acc * kPinv28) & kM28;
        acc += (uint64_t)m[k] * p28(0);
        acc >>= 28;
    }
    Fp28 r;
#pragma unroll
    for (int k = 14; k < 27; ++k) {
        uint64_t e = 0, o = 0, q = 0;
#pragma unroll
        for (int i = k - 13; i < 14; ++i) {
            if (i & 1) o += (uint64_t)a.v[i] * b.v[k - i];
            else e += (uint64_t)a.v[i] * b.v[k - i];
        }
#pragma unroll
        for (int i = k - 13; i < 14; ++i) q += (uint64_t)m[i] * p28(k - i);
        acc += e + o + q;
        r.v[k - 14] = (uint32_t)acc & kM28;
        acc >>= 28;
    }
    r.v[13] = (uint32_t)acc;
    fp28_reduce(r);
    return r;
}

// squaring: cross products once, doubled, plus the diagonal
template <bool kReduce = true>
__device__ __forceinline__ Fp28 fp28_sqr(const Fp28 &a)
{
    uint64_t acc = 0;
    uint32_t m[14];
#pragma unroll
    for (int k = 0; k < 14; ++k) {
        uint64_t cross = 0, q = 0, q2 = 0;
#pragma unroll
        for (int i = 0; 2 * i < k; ++i) cross += (uint64_t)a.v[i] * a.v[k - i];
        if ((k & 1) == 0) q2 = (uint64_t)a.v[k / 2] * a.v[k / 2];
#pragma unroll
        for (int i = 0; i < k; ++i) q += (uint64_t)m[i] * p28(k - i);
        acc += (cross << 1) + q2 + q;
        m[k] = ((uint32_t)acc * kPinv28) & kM28;
        acc += (uint64_t)m[k] * p28(0);
        acc >>= 28;
    }
    Fp28 r;
#pragma unroll
    for (int k = 14; k < 27; ++k) {
        uint64_t cross = 0, q = 0, q2 = 0;
#pragma unroll
        for (int i = k - 13; 2 * i < k; ++i) cross += (uint64_t)a.v[i] * a.v[k - i];
        if ((k & 1) == 0) q2 = (uint64_t)a.v[k / 2] * a.v[k / 2];
#pragma unroll
        for (int i = k - 13; i < 14; ++i) q += (uint64_t)m[i] * p28(k - i);
        acc += (cross << 1) + q2 + q;
        r.v[k - 14] = (uint32_t)acc & kM28;
        acc >>= 28;
    }
    r.v[13] = (uint32_t)acc;
    if (kReduce) fp28_reduce(r);
    return r;
}

__device__ __forceinline__ Fp28 fp28_add(const Fp28 &a, const Fp28 &b)
{
    Fp28 r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        uint32_t s = a.v[i] + b.v[i] + c;
        c = s >> 28;
        r.v[i] = s & kM28;
    }
    fp28_reduce(r);
    return r;
}

__device__ __forceinline__ Fp28 fp28_sub(const Fp28 &a, const Fp28 &b)
{
    Fp28 r;
    int32_t bw = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        int32_t t = (int32_t)a.v[i] - (int32_t)b.v[i] - bw;
        bw = (int32_t)((uint32_t)t >> 31);
        r.v[i] = (uint32_t)t & kM28;
    }
    // borrowed: add p back
    const uint32_t mask = 0u - (uint32_t)bw;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        uint32_t s = r.v[i] + (p28(i) & mask) + c;
        c = s >> 28;
        r.v[i] = s & kM28;
    }
    return r;
}

__device__ __forceinline__ Fp28 fp28_dbl(const Fp28 &a) { return fp28_add(a, a); }
__device__ __forceinline__ Fp28 fp28_neg(const Fp28 &a) { return fp28_sub(fp28_zero(), a); }

// ---- repacking between 12 x 32-bit and 14 x 28-bit limbs (same integer)

__device__ __forceinline__ Fp28 repack_to28(const Fp &x)
{
    Fp28 r;
#pragma unroll
    for (int j = 0; j < 14; ++j) {
        const int bit = 28 * j, q = bit >> 5, s = bit & 31;
        uint32_t lo = x.v[q];
        uint32_t hi = q + 1 < 12 ? x.v[q + 1] : 0u;
        uint64_t pair = (uint64_t)lo | ((uint64_t)hi << 32);
        r.v[j] = (uint32_t)(pair >> s) & kM28;
    }
    return r;
}

__device__ __forceinline__ Fp repack_to32(const Fp28 &x)
{
    Fp r;
#pragma unroll
    for (int w = 0; w < 12; ++w) {
        const int bit = 32 * w, j = bit / 28, s = bit % 28;
        uint32_t lo = x.v[j] >> s;
        uint32_t hi = j + 1 < 14 ? (x.v[j + 1] << (28 - s)) : 0u;
        r.v[w] = lo | hi;
    }
    return r;
}

__device__ __forceinline__ Fp28 fp28_from_std(const Fp &x) { return fp28_mul(repack_to28(x), fp28_const(0)); }
__device__ __forceinline__ Fp fp28_to_std(const Fp28 &x) { return repack_to32(fp28_mul(x, fp28_const(1))); }

// ---- points

struct Affine28 {
    Fp28 x, y;
    __device__ __forceinline__ bool is_inf() const { return fp28_is_zero(x) && fp28_is_zero(y); }
};

struct Xyzz28 {
    Fp28 x, y, zz, zzz;
    __device__ __forceinline__ bool is_inf() const { return fp28_is_zero(zz); }
};

__device__ __forceinline__ Xyzz28 xyzz28_inf()
{
    Xyzz28 r;
    r.x = fp28_one();
    r.y = fp28_one();
    r.zz = fp28_zero();
    r.zzz = fp28_zero();
    return r;
}

__device__ __forceinline__ Xyzz28 xyzz28_from_affine(const Affine28 &a)
{
    if (a.is_inf()) return xyzz28_inf();
    Xyzz28 r;
    r.x = a.x;
    r.y = a.y;
    r.zz = fp28_one();
    r.zzz = fp28_one();
    return r;
}

// mdbl-2008-s-1 on an affine point (not the identity)
__device__ __forceinline__ Xyzz28 xyzz28_dbl_affine(const Affine28 &q)
{
    if (fp28_is_zero(q.y)) return xyzz28_inf();
    Fp28 u = fp28_dbl(q.y);
    Fp28 v = fp28_sqr(u);
    Fp28 w = fp28_mul(u, v);
    Fp28 s = fp28_mul(q.x, v);
    Fp28 x2 = fp28_sqr(q.x);
    Fp28 m = fp28_add(x2, fp28_dbl(x2));
    Xyzz28 r;
    r.x = fp28_sub(fp28_sqr(m), fp28_dbl(s));
    r.y = fp28_sub(fp28_mul(m, fp28_sub(s, r.x)), fp28_mul(w, q.y));
    r.zz = v;
    r.zzz = w;
    return r;
}

// madd-2008-s: p + q with q affine (identity encoded as 0, 0)
__device__ __forceinline__ Xyzz28 xyzz28_add_affine(const Xyzz28 &p, const Affine28 &q)
{
    if (q.is_inf()) return p;
    if (p.is_inf()) return xyzz28_from_affine(q);
    Fp28 u2 = fp28_mul(q.x, p.zz);
    Fp28 s2 = fp28_mul(q.y, p.zzz);
    Fp28 pp_ = fp28_sub(u2, p.x);
    Fp28 r_ = fp28_sub(s2, p.y);
    if (fp28_is_zero(pp_)) {
        if (fp28_is_zero(r_)) return xyzz28_dbl_affine(q);
        return xyzz28_inf();
    }
    Fp28 pp = fp28_sqr(pp_);
    Fp28 ppp = fp28_mul(pp_, pp);
    Fp28 qq = fp28_mul(p.x, pp);
    Xyzz28 r;
    r.x = fp28_sub(fp28_sub(fp28_sqr(r_), ppp), fp28_dbl(qq));
    r.y = fp28_sub(fp28_mul(r_, fp28_sub(qq, r.x)), fp28_mul(p.y, ppp));
    r.zz = fp28_mul(p.zz, pp);
    r.zzz = fp28_mul(p.zzz, ppp);
    return r;
}

// ---- loose arithmetic for the accumulation loop.  Values lie in [0, 2p)
// instead of [0, p): a Montgomery product of two operands below 2p is
// a b / R + m p / R < 4 p^2 / 2^392 + p < 2p (p < 2^381), so products skip
// the final subtraction (56 of the ~583 instructions of a product); sums and
// differences are brought back below 2p with one correction by 2p, as the
// exact forms are with p.  A loose value is zero mod p iff it is 0 or p.

__device__ __forceinline__ uint32_t p2_28(int i)   // 2p
{
    constexpr uint32_t P2[14] = {0xfff5556u, 0xfdfffffu, 0x7ffff73u, 0xfffd62au, 0xc483d57u, 0x41ed61eu, 0xece61a5u,
                                 0xe70a257u, 0x8ee9709u, 0x9759aecu, 0x74f6c86u, 0xcd34963u, 0x3d472ffu, 0x0034022u};
    return P2[i];
}

// x - 2p if x >= 2p (x < 4p)
__device__ __forceinline__ void fp28_reduce2(Fp28 &x)
{
    uint32_t d[14];
    int32_t bw = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        int32_t t = (int32_t)x.v[i] - (int32_t)p2_28(i) - bw;
        bw = (int32_t)((uint32_t)t >> 31);
        d[i] = (uint32_t)t & kM28;
    }
#pragma unroll
    for (int i = 0; i < 14; ++i) x.v[i] = bw ? x.v[i] : d[i];
}

__device__ __forceinline__ Fp28 fp28_add_lz(const Fp28 &a, const Fp28 &b)
{
    Fp28 r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        uint32_t s = a.v[i] + b.v[i] + c;
        c = s >> 28;
        r.v[i] = s & kM28;
    }
    fp28_reduce2(r);
    return r;
}

__device__ __forceinline__ Fp28 fp28_sub_lz(const Fp28 &a, const Fp28 &b)
{
    Fp28 r;
    int32_t bw = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        int32_t t = (int32_t)a.v[i] - (int32_t)b.v[i] - bw;
        bw = (int32_t)((uint32_t)t >> 31);
        r.v[i] = (uint32_t)t & kM28;
    }
    // borrowed: add 2p back
    const uint32_t mask = 0u - (uint32_t)bw;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        uint32_t s = r.v[i] + (p2_28(i) & mask) + c;
        c = s >> 28;
        r.v[i] = s & kM28;
    }
    return r;
}

__device__ __forceinline__ bool fp28_is_zero_lz(const Fp28 &a)
{
    uint32_t z = 0, e = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        z |= a.v[i];
        e |= a.v[i] ^ p28(i);
    }
    return z == 0 || e == 0;
}

// ---------------------------------------------------------------------------
// Lazy forms for the accumulation's madd: limb-wise sums and differences with
// no carries and no conditional correction, for values that only feed a
// product (whose 64-bit accumulators take operand limbs below 2^29.6 here:
// each row then adds at most 2^57.6 + 2^57 + 2^56 < 2^58.5, fourteen rows
// stay below 2^62.3) or that are normalized and reduced at once.

// k p with 2^s lent to every limb below the top (the top limb gives what
// the limb below it borrowed), so that a subtrahend with limbs below 2^s
// never drives a limb negative: 2p and 4p with 2^28 lent, 8p with 2^30
__device__ __forceinline__ uint32_t p2x28_lent(int i)
{
    constexpr uint32_t L[14] = {0x1fff5556u, 0x1fdffffeu, 0x17ffff72u, 0x1fffd629u, 0x1c483d56u,
                                0x141ed61du, 0x1ece61a4u, 0x1e70a256u, 0x18ee9708u, 0x19759aebu,
                                0x174f6c85u, 0x1cd34962u, 0x13d472feu, 0x00034021u};
    return L[i];
}
__device__ __forceinline__ uint32_t p4x28_lent(int i)
{
    constexpr uint32_t L[14] = {0x1ffeaaacu, 0x1fbffffeu, 0x1ffffee6u, 0x1fffac53u, 0x18907aaeu,
                                0x183dac3cu, 0x1d9cc349u, 0x1ce144aeu, 0x11dd2e12u, 0x12eb35d8u,
                                0x1e9ed90cu, 0x19a692c5u, 0x17a8e5feu, 0x00068043u};
    return L[i];
}
__device__ __forceinline__ uint32_t p8x30_lent(int i)
{
    constexpr uint32_t L[14] = {0x4ffd5558u, 0x4f7ffffbu, 0x4ffffdcbu, 0x4fff58a5u, 0x4120f55bu,
                                0x407b5877u, 0x4b398691u, 0x49c2895bu, 0x43ba5c23u, 0x45d66baeu,
                                0x4d3db216u, 0x434d2589u, 0x4f51cbfbu, 0x000d0084u};
    return L[i];
}

// 2p - y for canonical y (top limb below that of 2p minus one): the
// negation, below 2p with limbs below 2^29, as a product operand
__device__ __forceinline__ Fp28 fp28_lazy_neg2p(const Fp28 &y)
{
    Fp28 r;
#pragma unroll
    for (int i = 0; i < 14; ++i) r.v[i] = p2x28_lent(i) - y.v[i];
    return r;
}

// a - b + 4p for a, b loose (below 2p, limbs below 2^28): below 6p, limbs
// below 2^29.6, as a product operand
__device__ __forceinline__ Fp28 fp28_lazy_sub4p(const Fp28 &a, const Fp28 &b)
{
    Fp28 r;
#pragma unroll
    for (int i = 0; i < 14; ++i) r.v[i] = a.v[i] + p4x28_lent(i) - b.v[i];
    return r;
}

// u - x for loose u, x is 0 mod p only if it is 0, p or -p, whose low 28
// bits are 0, p_0 and 2^28 - p_0: any other low difference rules it out
// (the exact test runs for about three pairs in 2^28)
__device__ __forceinline__ bool fp28_may_equal(const Fp28 &u, const Fp28 &x)
{
    const uint32_t d = (u.v[0] - x.v[0]) & kM28;
    return d == 0 || d == 0xfffaaabu || d == 0x0005555u;
}

// normalized x (limbs below 2^28, value below 2^392) -> the same class below
// 2p: x - q p with q = floor(x13 M / 2^48), M = floor(2^48 / (p13 + 1)), p13
// the top limb of p.  x13 M / 2^48 is below x / p and short of it by at
// most (x / p) / p13 + 2^-16, so q is floor(x / p) or one less for x / p
// up to ~1000.
__device__ __forceinline__ void fp28_reduce_top(Fp28 &a)
{
    const uint32_t q = (uint32_t)(((uint64_t)a.v[13] * 0x9d8307deull) >> 48);
    int64_t carry = 0;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        const int64_t t = (int64_t)a.v[i] - (int64_t)q * (int64_t)p28(i) + carry;
        a.v[i] = i < 13 ? (uint32_t)t & kM28 : (uint32_t)t;
        carry = t >> 28;   // arithmetic: the borrow into the next limb
    }
}

// X3 = r2 - ppp - 2 qq for loose r2, ppp, qq: formed limb-wise with 8p
// added (below 10p, limbs below 2^31), then carried and reduced by a
// quotient from the top limb: loose (below 2p, limbs below 2^28)
__device__ __forceinline__ Fp28 fp28_x3_lz(const Fp28 &r2, const Fp28 &ppp, const Fp28 &qq)
{
    Fp28 r;
#pragma unroll
    for (int i = 0; i < 14; ++i) r.v[i] = r2.v[i] + p8x30_lent(i) - ppp.v[i] - 2 * qq.v[i];
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 13; ++i) {
        const uint32_t s = r.v[i] + c;
        c = s >> 28;
        r.v[i] = s & kM28;
    }
    r.v[13] += c;
    fp28_reduce_top(r);
    return r;
}

// madd-2008-s on loose coordinates: p (accumulator, loose, identity when
// zz == 0 exactly) + q or -q (neg; q affine, reduced).  The identity is
// only ever made by xyzz28_inf, and zz of a sum is zz * PP with PP nonzero,
// so the exact zero test on zz stays valid.  The negated y enters the one
// product that takes it as 2p - y, unreduced; X3 is formed lazily and
// reduced once, and Y3's two differences stay lazy inside its two-product
// reduction.  The rare exceptional paths take -q exactly.
__device__ __forceinline__ Xyzz28 xyzz28_add_signed_lz(const Xyzz28 &p, const Affine28 &q, bool neg)
{
    if (q.is_inf()) return p;
    if (p.is_inf()) {
        Affine28 c = q;
        if (neg) c.y = fp28_neg(q.y);
        return xyzz28_from_affine(c);
    }
    const Fp28 ny = fp28_lazy_neg2p(q.y);
    Fp28 qy;
#pragma unroll
    for (int i = 0; i < 14; ++i) qy.v[i] = neg ? ny.v[i] : q.y.v[i];
    Fp28 u2 = fp28_mul<false>(q.x, p.zz);
    Fp28 s2 = fp28_mul<false>(qy, p.zzz);
    if (fp28_may_equal(u2, p.x) && fp28_is_zero_lz(fp28_sub_lz(u2, p.x))) {
        if (fp28_is_zero_lz(fp28_sub_lz(s2, p.y))) {
            Affine28 c = q;
            if (neg) c.y = fp28_neg(q.y);
            return xyzz28_dbl_affine(c);
        }
        return xyzz28_inf();
    }
    // PP_ and R stay lazy (plus 4p: below 6p, limbs below 2^29.6)
    const Fp28 pp_ = fp28_lazy_sub4p(u2, p.x);
    const Fp28 r_ = fp28_lazy_sub4p(s2, p.y);
    Fp28 pp = fp28_sqr<false>(pp_);
    Fp28 ppp = fp28_mul<false>(pp_, pp);
    Fp28 qq = fp28_mul<false>(p.x, pp);
    Xyzz28 r;
    r.x = fp28_x3_lz(fp28_sqr<false>(r_), ppp, qq);
    // Y3 = R (Q - X3) - Y1 PPP = R (Q - X3 + 4p) + (4p - Y1) PPP
    r.y = fp28_mul2_lz(r_, fp28_lazy_sub4p(qq, r.x), fp28_lazy_sub4p(fp28_zero(), p.y), ppp);
    r.zz = fp28_mul<false>(p.zz, pp);
    r.zzz = fp28_mul<false>(p.zzz, ppp);
    return r;
}

__device__ __forceinline__ Affine28 load_affine28(const uint32_t *p)
{
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    Affine28 a;
    uint32_t t[28];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        uint4 u = q[i];
        t[4 * i] = u.x;
        t[4 * i + 1] = u.y;
        t[4 * i + 2] = u.z;
        t[4 * i + 3] = u.w;
    }
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        a.x.v[i] = t[i];
        a.y.v[i] = t[14 + i];
    }
    return a;
}

__device__ __forceinline__ void store_affine28(uint32_t *p, const Affine28 &a)
{
    uint4 *q = reinterpret_cast<uint4 *>(p);
    uint32_t t[28];
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        t[i] = a.x.v[i];
        t[14 + i] = a.y.v[i];
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) q[i] = make_uint4(t[4 * i], t[4 * i + 1], t[4 * i + 2], t[4 * i + 3]);
}

__device__ __forceinline__ Affine28 affine28_from_std(const Affine &a)
{
    Affine28 r;
    r.x = fp28_from_std(a.x);
    r.y = fp28_from_std(a.y);
    return r;
}

__device__ __forceinline__ Xyzz xyzz28_to_std(const Xyzz28 &p)
{
    if (p.is_inf()) return Xyzz::inf();
    Xyzz r;
    r.x = fp28_to_std(p.x);
    r.y = fp28_to_std(p.y);
    r.zz = fp28_to_std(p.zz);
    r.zzz = fp28_to_std(p.zzz);
    return r;
}

// ---- loose XYZZ points in memory and the general loose addition, for the
// bucket reduction that follows the accumulation: partial sums stay in
// radix 2^28 (loose, [0, 2p)) until the last group pass converts them.

constexpr uint32_t kXyzz28Words = 56;   // 4 x 14 limbs, 224 B (14 x 16 B)

__device__ __forceinline__ Xyzz28 load_xyzz28(const uint32_t *p)
{
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    uint32_t t[56];
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        const uint4 u = q[i];
        t[4 * i] = u.x;
        t[4 * i + 1] = u.y;
        t[4 * i + 2] = u.z;
        t[4 * i + 3] = u.w;
    }
    Xyzz28 a;
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        a.x.v[i] = t[i];
        a.y.v[i] = t[14 + i];
        a.zz.v[i] = t[28 + i];
        a.zzz.v[i] = t[42 + i];
    }
    return a;
}

__device__ __forceinline__ void store_xyzz28(uint32_t *p, const Xyzz28 &a)
{
    uint32_t t[56];
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        t[i] = a.x.v[i];
        t[14 + i] = a.y.v[i];
        t[28 + i] = a.zz.v[i];
        t[42 + i] = a.zzz.v[i];
    }
    uint4 *q = reinterpret_cast<uint4 *>(p);
#pragma unroll
    for (int i = 0; i < 14; ++i) q[i] = make_uint4(t[4 * i], t[4 * i + 1], t[4 * i + 2], t[4 * i + 3]);
}

// 32-bit form -> loose radix 2^28 (one product per coordinate); the
// identity keeps zz = 0 exactly
__device__ __forceinline__ Xyzz28 xyzz28_from_std_lz(const Xyzz &p)
{
    if (p.is_inf()) return xyzz28_inf();
    Xyzz28 r;
    r.x = fp28_mul<false>(repack_to28(p.x), fp28_const(0));
    r.y = fp28_mul<false>(repack_to28(p.y), fp28_const(0));
    r.zz = fp28_mul<false>(repack_to28(p.zz), fp28_const(0));
    r.zzz = fp28_mul<false>(repack_to28(p.zzz), fp28_const(0));
    return r;
}

// dbl-2008-s-1 on loose coordinates
__device__ __forceinline__ Xyzz28 xyzz28_dbl_lz(const Xyzz28 &p)
{
    if (p.is_inf() || fp28_is_zero_lz(p.y)) return xyzz28_inf();
    Fp28 u = fp28_add_lz(p.y, p.y);
    Fp28 v = fp28_sqr<false>(u);
    Fp28 w = fp28_mul<false>(u, v);
    Fp28 s = fp28_mul<false>(p.x, v);
    Fp28 x2 = fp28_sqr<false>(p.x);
    Fp28 m = fp28_add_lz(x2, fp28_add_lz(x2, x2));
    Xyzz28 r;
    r.x = fp28_sub_lz(fp28_sqr<false>(m), fp28_add_lz(s, s));
    r.y = fp28_sub_lz(fp28_mul<false>(m, fp28_sub_lz(s, r.x)), fp28_mul<false>(w, p.y));
    r.zz = fp28_mul<false>(v, p.zz);
    r.zzz = fp28_mul<false>(w, p.zzz);
    return r;
}

// add-2008-s on loose coordinates (the identity only as zz = 0 exactly: zz
// of a sum is zz1 zz2 PP and of a double v zz, both nonzero mod p)
__device__ __forceinline__ Xyzz28 xyzz28_add_lz(const Xyzz28 &p, const Xyzz28 &q)
{
    if (q.is_inf()) return p;
    if (p.is_inf()) return q;
    Fp28 u1 = fp28_mul<false>(p.x, q.zz);
    Fp28 u2 = fp28_mul<false>(q.x, p.zz);
    Fp28 s1 = fp28_mul<false>(p.y, q.zzz);
    Fp28 s2 = fp28_mul<false>(q.y, p.zzz);
    if (fp28_may_equal(u2, u1) && fp28_is_zero_lz(fp28_sub_lz(u2, u1))) {
        if (fp28_is_zero_lz(fp28_sub_lz(s2, s1))) return xyzz28_dbl_lz(p);
        return xyzz28_inf();
    }
    const Fp28 pp_ = fp28_lazy_sub4p(u2, u1);
    const Fp28 r_ = fp28_lazy_sub4p(s2, s1);
    Fp28 pp = fp28_sqr<false>(pp_);
    Fp28 ppp = fp28_mul<false>(pp_, pp);
    Fp28 qq = fp28_mul<false>(u1, pp);
    Xyzz28 r;
    r.x = fp28_x3_lz(fp28_sqr<false>(r_), ppp, qq);
    // Y3 = R (Q - X3) - S1 PPP = R (Q - X3 + 4p) + (4p - S1) PPP, one
    // reduction, the differences lazy
    r.y = fp28_mul2_lz(r_, fp28_lazy_sub4p(qq, r.x), fp28_lazy_sub4p(fp28_zero(), s1), ppp);
    r.zz = fp28_mul<false>(fp28_mul<false>(p.zz, q.zz), pp);
    r.zzz = fp28_mul<false>(fp28_mul<false>(p.zzz, q.zzz), ppp);
    return r;
}

// ------------------------------------------------- four lanes per addition
//
// A pass with too few partials to give every SIMD a wave is bound by the
// latency of one lane's chain of additions, not by issue.  Four lanes (a DPP
// quad) then share one addition: add-2008-s has its fourteen products in
// four dependent steps of at most four, so each lane runs one product per
// step and the quad broadcasts the four results.  Every lane of a quad holds
// the same operands and results, so the quad takes every branch together,
// and each product is the same Montgomery product of the same operands as
// in xyzz28_add_lz except Y3, which the one-lane form folds into one
// reduction: the sum is the same point (a representative of the same
// coordinates mod p) as the one-lane form's.

// lane `kLane` of each quad -> all four lanes of the quad
template <int kLane>
__device__ __forceinline__ Fp28 fp28_quad_bcast(const Fp28 &a)
{
    Fp28 r;
#pragma unroll
    for (int i = 0; i < 14; ++i)
        r.v[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)a.v[i], kLane * 0x55, 0xf, 0xf, false);
    return r;
}

// operand of quad lane `sub`: masks and ORs per limb (a chain of selects
// is turned into a stack table indexed by `sub`)
__device__ __forceinline__ Fp28 fp28_sel4(uint32_t sub, const Fp28 &a0, const Fp28 &a1, const Fp28 &a2,
                                          const Fp28 &a3)
{
    const uint32_t m0 = 0u - (uint32_t)(sub == 0), m1 = 0u - (uint32_t)(sub == 1);
    const uint32_t m2 = 0u - (uint32_t)(sub == 2), m3 = 0u - (uint32_t)(sub == 3);
    Fp28 r;
#pragma unroll
    for (int i = 0; i < 14; ++i) r.v[i] = (a0.v[i] & m0) | (a1.v[i] & m1) | (a2.v[i] & m2) | (a3.v[i] & m3);
    return r;
}

// xyzz28_add_lz on a quad; `sub` is the lane's place in its quad
__device__ __forceinline__ Xyzz28 xyzz28_add_quad(const Xyzz28 &p, const Xyzz28 &q, uint32_t sub)
{
    if (q.is_inf()) return p;
    if (p.is_inf()) return q;
    // u1 = x1 zz2, u2 = x2 zz1, s1 = y1 zzz2, s2 = y2 zzz1
    Fp28 t = fp28_mul<false>(fp28_sel4(sub, p.x, q.x, p.y, q.y), fp28_sel4(sub, q.zz, p.zz, q.zzz, p.zzz));
    const Fp28 u1 = fp28_quad_bcast<0>(t), s1 = fp28_quad_bcast<2>(t);
    const Fp28 u2 = fp28_quad_bcast<1>(t), s2 = fp28_quad_bcast<3>(t);
    if (fp28_may_equal(u2, u1) && fp28_is_zero_lz(fp28_sub_lz(u2, u1))) {
        if (fp28_is_zero_lz(fp28_sub_lz(s2, s1))) return xyzz28_dbl_lz(p);
        return xyzz28_inf();
    }
    const Fp28 pp_ = fp28_lazy_sub4p(u2, u1);
    const Fp28 r_ = fp28_lazy_sub4p(s2, s1);
    // PP = P^2, R^2, zz1 zz2, zzz1 zzz2
    t = fp28_mul<false>(fp28_sel4(sub, pp_, r_, p.zz, p.zzz), fp28_sel4(sub, pp_, r_, q.zz, q.zzz));
    const Fp28 pp = fp28_quad_bcast<0>(t), r2 = fp28_quad_bcast<1>(t);
    const Fp28 zz12 = fp28_quad_bcast<2>(t), zzz12 = fp28_quad_bcast<3>(t);
    // PPP = P PP, Q = u1 PP, zz3 = zz1 zz2 PP (lane 3 repeats lane 2)
    t = fp28_mul<false>(fp28_sel4(sub, pp_, u1, zz12, zz12), pp);
    const Fp28 ppp = fp28_quad_bcast<0>(t), qq = fp28_quad_bcast<1>(t);
    Xyzz28 r;
    r.zz = fp28_quad_bcast<2>(t);
    r.x = fp28_x3_lz(r2, ppp, qq);
    // R (Q - x3), s1 PPP, zzz3 = zzz1 zzz2 PPP (Q - x3 lazy, plus 4p)
    t = fp28_mul<false>(fp28_sel4(sub, r_, s1, zzz12, zzz12), fp28_sel4(sub, fp28_lazy_sub4p(qq, r.x), ppp, ppp, ppp));
    r.y = fp28_sub_lz(fp28_quad_bcast<0>(t), fp28_quad_bcast<1>(t));
    r.zzz = fp28_quad_bcast<2>(t);
    return r;
}

// xyzz28_add_signed_lz on a quad: its ten products in four steps
// (u2 s2 | PP R^2 | PPP Q zz3 | R(Q-X3) y1PPP zzz3)
__device__ __forceinline__ Xyzz28 xyzz28_add_affine_quad(const Xyzz28 &p, const Affine28 &q, bool neg, uint32_t sub)
{
    if (q.is_inf()) return p;
    if (p.is_inf()) {
        Affine28 c = q;
        if (neg) c.y = fp28_neg(q.y);
        return xyzz28_from_affine(c);
    }
    const Fp28 ny = fp28_lazy_neg2p(q.y);
    Fp28 qy;
#pragma unroll
    for (int i = 0; i < 14; ++i) qy.v[i] = neg ? ny.v[i] : q.y.v[i];
    Fp28 t = fp28_mul<false>(fp28_sel4(sub, q.x, qy, q.x, qy), fp28_sel4(sub, p.zz, p.zzz, p.zz, p.zzz));
    const Fp28 u2 = fp28_quad_bcast<0>(t), s2 = fp28_quad_bcast<1>(t);
    if (fp28_may_equal(u2, p.x) && fp28_is_zero_lz(fp28_sub_lz(u2, p.x))) {
        if (fp28_is_zero_lz(fp28_sub_lz(s2, p.y))) {
            Affine28 c = q;
            if (neg) c.y = fp28_neg(q.y);
            return xyzz28_dbl_affine(c);
        }
        return xyzz28_inf();
    }
    const Fp28 pp_ = fp28_lazy_sub4p(u2, p.x);
    const Fp28 r_ = fp28_lazy_sub4p(s2, p.y);
    const Fp28 sq = fp28_sel4(sub, pp_, r_, pp_, r_);
    t = fp28_mul<false>(sq, sq);
    const Fp28 pp = fp28_quad_bcast<0>(t), r2 = fp28_quad_bcast<1>(t);
    t = fp28_mul<false>(fp28_sel4(sub, pp_, p.x, p.zz, p.zz), pp);
    const Fp28 ppp = fp28_quad_bcast<0>(t), qq = fp28_quad_bcast<1>(t);
    Xyzz28 r;
    r.zz = fp28_quad_bcast<2>(t);
    r.x = fp28_x3_lz(r2, ppp, qq);
    t = fp28_mul<false>(fp28_sel4(sub, r_, p.y, p.zzz, p.zzz), fp28_sel4(sub, fp28_lazy_sub4p(qq, r.x), ppp, ppp, ppp));
    r.y = fp28_sub_lz(fp28_quad_bcast<0>(t), fp28_quad_bcast<1>(t));
    r.zzz = fp28_quad_bcast<2>(t);
    return r;
}

}  // namespace mh
