// BLS12-381 Fr / Fp Montgomery arithmetic for gfx950.
//
// Limbs are 32-bit, little-endian, Montgomery form (Fr: R = 2^256, 8 limbs;
// Fp: R = 2^384, 12 limbs), so a device element has exactly the bytes of the
// reference ABI types RawFr (4 x u64) and the coordinates of RawG1Affine
// (6 x u64) (midnight-sppark-msm/src/lib.rs:17-36).  Every operation returns
// a fully reduced value in [0, p): equal field elements have equal limbs,
// which is what makes NTT outputs limb-identical to the CPU prover.
//
// The product is CIOS Montgomery multiplication (Koc, Acar, Kaliski 1996)
// arranged for v_mad_u64_u32: each row forms the 64-bit partial products
// a_j*b_i + t_j independently (one v_mad_u64_u32 each) and then folds their
// high halves into the next word with one add-with-carry chain, so a row
// costs one multiply-add and one v_addc per limb.  Both moduli leave spare
// top bits (r < 2^255, p < 2^381), so the running value stays below 2p and
// needs no extra overflow word.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mh {

struct FrCfg {
    static constexpr int N = 8;
    static constexpr uint32_t INV = 0xffffffffu;   // -r^-1 mod 2^32
    __device__ static constexpr uint32_t p(int i)
    {
        constexpr uint32_t P[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                                   0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
        return P[i];
    }
    __device__ static constexpr uint32_t one(int i)   // R mod r
    {
        constexpr uint32_t O[8] = {0xfffffffeu, 0x00000001u, 0x00034802u, 0x5884b7fau,
                                   0xecbc4ff5u, 0x998c4fefu, 0xacc5056fu, 0x1824b159u};
        return O[i];
    }
};

struct FpCfg {
    static constexpr int N = 12;
    static constexpr uint32_t INV = 0xfffcfffdu;   // -p^-1 mod 2^32
    __device__ static constexpr uint32_t p(int i)
    {
        constexpr uint32_t P[12] = {0xffffaaabu, 0xb9feffffu, 0xb153ffffu, 0x1eabfffeu,
                                    0xf6b0f624u, 0x6730d2a0u, 0xf38512bfu, 0x64774b84u,
                                    0x434bacd7u, 0x4b1ba7b6u, 0x397fe69au, 0x1a0111eau};
        return P[i];
    }
    __device__ static constexpr uint32_t one(int i)   // R mod p
    {
        constexpr uint32_t O[12] = {0x0002fffdu, 0x76090000u, 0xc40c0002u, 0xebf4000bu,
                                    0x53c758bau, 0x5f489857u, 0x70525745u, 0x77ce5853u,
                                    0xa256ec6du, 0x5c071a97u, 0xfa80e493u, 0x15f65ec3u};
        return O[i];
    }
};

__device__ __forceinline__ uint32_t addc(uint32_t a, uint32_t b, uint32_t cin, uint32_t &cout)
{
    uint32_t c;
    uint32_t r = __builtin_addc(a, b, cin, &c);
    cout = c;
    return r;
}

__device__ __forceinline__ uint32_t subb(uint32_t a, uint32_t b, uint32_t bin, uint32_t &bout)
{
    uint32_t c;
    uint32_t r = __builtin_subc(a, b, bin, &c);
    bout = c;
    return r;
}

template <class C>
struct Fe {
    static constexpr int N = C::N;
    uint32_t v[N];

    __device__ __forceinline__ static Fe zero()
    {
        Fe r;
#pragma unroll
        for (int i = 0; i < N; ++i) r.v[i] = 0;
        return r;
    }
    __device__ __forceinline__ static Fe one()
    {
        Fe r;
#pragma unroll
        for (int i = 0; i < N; ++i) r.v[i] = C::one(i);
        return r;
    }
    __device__ __forceinline__ bool is_zero() const
    {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) o |= v[i];
        return o == 0;
    }
    __device__ __forceinline__ bool eq(const Fe &b) const
    {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) o |= v[i] ^ b.v[i];
        return o == 0;
    }
};

using Fr = Fe<FrCfg>;
using Fp = Fe<FpCfg>;

// r = x - p if x >= p else x  (x < 2p)
template <class C>
__device__ __forceinline__ void reduce_once(uint32_t *x)
{
    constexpr int N = C::N;
    uint32_t d[N], b = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = subb(x[i], C::p(i), b, b);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = b ? x[i] : d[i];
}

template <class C>
__device__ __forceinline__ Fe<C> add(const Fe<C> &a, const Fe<C> &b)
{
    Fe<C> r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < C::N; ++i) r.v[i] = addc(a.v[i], b.v[i], c, c);
    // the sum is < 2p < 2^(32N): no carry out
    reduce_once<C>(r.v);
    return r;
}

template <class C>
__device__ __forceinline__ Fe<C> sub(const Fe<C> &a, const Fe<C> &b)
{
    constexpr int N = C::N;
    Fe<C> r;
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) r.v[i] = subb(a.v[i], b.v[i], br, br);
    // add p back when the subtraction borrowed
    uint32_t mask = 0u - br, c = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) r.v[i] = addc(r.v[i], C::p(i) & mask, c, c);
    return r;
}

template <class C>
__device__ __forceinline__ Fe<C> dbl(const Fe<C> &a)
{
    return add<C>(a, a);
}

template <class C>
__device__ __forceinline__ Fe<C> neg(const Fe<C> &a)
{
    return sub<C>(Fe<C>::zero(), a);
}

template <class C>
__device__ __forceinline__ Fe<C> mul(const Fe<C> &a, const Fe<C> &b)
{
    constexpr int N = C::N;
    uint32_t t[N + 1];
#pragma unroll
    for (int j = 0; j <= N; ++j) t[j] = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        // t += a * b_i
        uint64_t pr[N];
#pragma unroll
        for (int j = 0; j < N; ++j) pr[j] = (uint64_t)a.v[j] * b.v[i] + t[j];
        uint32_t c = 0;
        t[0] = (uint32_t)pr[0];
#pragma unroll
        for (int j = 1; j < N; ++j) t[j] = addc((uint32_t)pr[j], (uint32_t)(pr[j - 1] >> 32), c, c);
        t[N] = (uint32_t)(pr[N - 1] >> 32) + c;
        // t = (t + m p) / 2^32
        uint32_t m = t[0] * C::INV;
        uint64_t q[N];
#pragma unroll
        for (int j = 0; j < N; ++j) q[j] = (uint64_t)m * C::p(j) + t[j];
        c = 0;
#pragma unroll
        for (int j = 1; j < N; ++j) t[j - 1] = addc((uint32_t)q[j], (uint32_t)(q[j - 1] >> 32), c, c);
        t[N - 1] = t[N] + (uint32_t)(q[N - 1] >> 32) + c;
        t[N] = 0;
    }
    Fe<C> r;
#pragma unroll
    for (int j = 0; j < N; ++j) r.v[j] = t[j];
    reduce_once<C>(r.v);
    return r;
}

template <class C>
__device__ __forceinline__ Fe<C> sqr(const Fe<C> &a)
{
    return mul<C>(a, a);
}

// Operators for readability in kernels
__device__ __forceinline__ Fr operator+(const Fr &a, const Fr &b) { return add<FrCfg>(a, b); }
__device__ __forceinline__ Fr operator-(const Fr &a, const Fr &b) { return sub<FrCfg>(a, b); }
__device__ __forceinline__ Fr operator*(const Fr &a, const Fr &b) { return mul<FrCfg>(a, b); }
__device__ __forceinline__ Fp operator+(const Fp &a, const Fp &b) { return add<FpCfg>(a, b); }
__device__ __forceinline__ Fp operator-(const Fp &a, const Fp &b) { return sub<FpCfg>(a, b); }
__device__ __forceinline__ Fp operator*(const Fp &a, const Fp &b) { return mul<FpCfg>(a, b); }

// Fr from a canonical 256-bit value is never needed on device; Montgomery
// to canonical (multiply by 1) is, for scalar digit extraction.
__device__ __forceinline__ Fr fr_from_mont(const Fr &a)
{
    Fr one_raw = Fr::zero();
    one_raw.v[0] = 1;
    return mul<FrCfg>(a, one_raw);
}

// Vector loads/stores of whole elements (32 B / 48 B)
__device__ __forceinline__ Fr load_fr(const Fr *p)
{
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    uint4 x = q[0], y = q[1];
    Fr r;
    r.v[0] = x.x; r.v[1] = x.y; r.v[2] = x.z; r.v[3] = x.w;
    r.v[4] = y.x; r.v[5] = y.y; r.v[6] = y.z; r.v[7] = y.w;
    return r;
}

__device__ __forceinline__ void store_fr(Fr *p, const Fr &a)
{
    uint4 *q = reinterpret_cast<uint4 *>(p);
    q[0] = make_uint4(a.v[0], a.v[1], a.v[2], a.v[3]);
    q[1] = make_uint4(a.v[4], a.v[5], a.v[6], a.v[7]);
}

__device__ __forceinline__ Fp load_fp(const uint32_t *p)
{
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    uint4 x = q[0], y = q[1], z = q[2];
    Fp r;
    r.v[0] = x.x; r.v[1] = x.y; r.v[2] = x.z; r.v[3] = x.w;
    r.v[4] = y.x; r.v[5] = y.y; r.v[6] = y.z; r.v[7] = y.w;
    r.v[8] = z.x; r.v[9] = z.y; r.v[10] = z.z; r.v[11] = z.w;
    return r;
}

__device__ __forceinline__ void store_fp(uint32_t *p, const Fp &a)
{
    uint4 *q = reinterpret_cast<uint4 *>(p);
    q[0] = make_uint4(a.v[0], a.v[1], a.v[2], a.v[3]);
    q[1] = make_uint4(a.v[4], a.v[5], a.v[6], a.v[7]);
    q[2] = make_uint4(a.v[8], a.v[9], a.v[10], a.v[11]);
}

}  // namespace mh
