// BLS12-381 G1 group law on device (y^2 = x^3 + 4 over Fp).
//
// Bucket sums live in XYZZ coordinates (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2),
// which give an 8M+2S mixed addition and 12M+2S general addition
// (Explicit-Formulas Database: madd-2008-s, add-2008-s, dbl-2008-s-1,
// mdbl-2008-s-1).  The identity has ZZ = ZZZ = 0.  Results leave the device
// in Jacobian form (X*ZZ^2, Y*ZZZ^2, ZZZ) -- the reference's RawG1Projective
// (blst p1) layout, where Z = 0 is the identity.  Affine inputs use the
// reference's RawG1Affine layout with the identity encoded as x = y = 0.
#pragma once
#include "fields.hpp"

namespace mh {

struct Affine {
    Fp x, y;
    __device__ __forceinline__ bool is_inf() const { return x.is_zero() && y.is_zero(); }
};

struct Xyzz {
    Fp x, y, zz, zzz;
    __device__ __forceinline__ static Xyzz inf()
    {
        Xyzz r;
        r.x = Fp::one();
        r.y = Fp::one();
        r.zz = Fp::zero();
        r.zzz = Fp::zero();
        return r;
    }
    __device__ __forceinline__ bool is_inf() const { return zz.is_zero(); }
};

struct Jacobian {
    Fp x, y, z;
};

__device__ __forceinline__ Xyzz xyzz_from_affine(const Affine &a)
{
    if (a.is_inf()) return Xyzz::inf();
    Xyzz r;
    r.x = a.x;
    r.y = a.y;
    r.zz = Fp::one();
    r.zzz = Fp::one();
    return r;
}

// dbl-2008-s-1 (a = 0)
__device__ __forceinline__ Xyzz xyzz_dbl(const Xyzz &p)
{
    if (p.is_inf() || p.y.is_zero()) return Xyzz::inf();
    Fp u = dbl<FpCfg>(p.y);
    Fp v = sqr<FpCfg>(u);
    Fp w = u * v;
    Fp s = p.x * v;
    Fp x2 = sqr<FpCfg>(p.x);
    Fp m = x2 + dbl<FpCfg>(x2);
    Xyzz r;
    r.x = sqr<FpCfg>(m) - dbl<FpCfg>(s);
    r.y = m * (s - r.x) - w * p.y;
    r.zz = v * p.zz;
    r.zzz = w * p.zzz;
    return r;
}

// mdbl-2008-s-1: double an affine point (not the identity)
__device__ __forceinline__ Xyzz xyzz_dbl_affine(const Affine &q)
{
    if (q.y.is_zero()) return Xyzz::inf();
    Fp u = dbl<FpCfg>(q.y);
    Fp v = sqr<FpCfg>(u);
    Fp w = u * v;
    Fp s = q.x * v;
    Fp x2 = sqr<FpCfg>(q.x);
    Fp m = x2 + dbl<FpCfg>(x2);
    Xyzz r;
    r.x = sqr<FpCfg>(m) - dbl<FpCfg>(s);
    r.y = m * (s - r.x) - w * q.y;
    r.zz = v;
    r.zzz = w;
    return r;
}

// madd-2008-s: p + q with q affine (identity encoded as 0,0)
__device__ __forceinline__ Xyzz xyzz_add_affine(const Xyzz &p, const Affine &q)
{
    if (q.is_inf()) return p;
    if (p.is_inf()) return xyzz_from_affine(q);
    Fp u2 = q.x * p.zz;
    Fp s2 = q.y * p.zzz;
    Fp pp_ = u2 - p.x;
    Fp r_ = s2 - p.y;
    if (pp_.is_zero()) {
        if (r_.is_zero()) return xyzz_dbl_affine(q);
        return Xyzz::inf();
    }
    Fp pp = sqr<FpCfg>(pp_);
    Fp ppp = pp_ * pp;
    Fp qq = p.x * pp;
    Xyzz r;
    r.x = sqr<FpCfg>(r_) - ppp - dbl<FpCfg>(qq);
    r.y = r_ * (qq - r.x) - p.y * ppp;
    r.zz = p.zz * pp;
    r.zzz = p.zzz * ppp;
    return r;
}

// add-2008-s: general XYZZ addition
__device__ __forceinline__ Xyzz xyzz_add(const Xyzz &p, const Xyzz &q)
{
    if (q.is_inf()) return p;
    if (p.is_inf()) return q;
    Fp u1 = p.x * q.zz;
    Fp u2 = q.x * p.zz;
    Fp s1 = p.y * q.zzz;
    Fp s2 = q.y * p.zzz;
    Fp pp_ = u2 - u1;
    Fp r_ = s2 - s1;
    if (pp_.is_zero()) {
        if (r_.is_zero()) return xyzz_dbl(p);
        return Xyzz::inf();
    }
    Fp pp = sqr<FpCfg>(pp_);
    Fp ppp = pp_ * pp;
    Fp qq = u1 * pp;
    Xyzz r;
    r.x = sqr<FpCfg>(r_) - ppp - dbl<FpCfg>(qq);
    r.y = r_ * (qq - r.x) - s1 * ppp;
    r.zz = p.zz * q.zz * pp;
    r.zzz = p.zzz * q.zzz * ppp;
    return r;
}

__device__ __forceinline__ Xyzz xyzz_neg(const Xyzz &p)
{
    Xyzz r = p;
    r.y = neg<FpCfg>(p.y);
    return r;
}

__device__ __forceinline__ Affine affine_neg(const Affine &a)
{
    Affine r = a;
    r.y = neg<FpCfg>(a.y);   // 0 stays 0 for the identity
    return r;
}

__device__ __forceinline__ Jacobian xyzz_to_jacobian(const Xyzz &p)
{
    Jacobian j;
    if (p.is_inf()) {
        j.x = Fp::one();
        j.y = Fp::one();
        j.z = Fp::zero();
        return j;
    }
    j.x = p.x * sqr<FpCfg>(p.zz);
    j.y = p.y * sqr<FpCfg>(p.zzz);
    j.z = p.zzz;
    return j;
}

// a^(p-2) = a^-1 (Fermat); 0 maps to 0
__device__ __forceinline__ Fp fp_inverse(const Fp &a)
{
    // p - 2, 32-bit limbs, little-endian
    constexpr uint32_t E[12] = {0xffffaaa9u, 0xb9feffffu, 0xb153ffffu, 0x1eabfffeu, 0xf6b0f624u, 0x6730d2a0u,
                                0xf38512bfu, 0x64774b84u, 0x434bacd7u, 0x4b1ba7b6u, 0x397fe69au, 0x1a0111eau};
    Fp r = Fp::one();
    for (int i = 380; i >= 0; --i) {
        r = sqr<FpCfg>(r);
        if ((E[i >> 5] >> (i & 31)) & 1u) r = r * a;
    }
    return r;
}

__device__ __forceinline__ Affine xyzz_to_affine(const Xyzz &p)
{
    Affine a;
    if (p.is_inf()) {
        a.x = Fp::zero();
        a.y = Fp::zero();
        return a;
    }
    Fp inv = fp_inverse(p.zz * p.zzz);
    a.x = p.x * (inv * p.zzz);   // X / ZZ
    a.y = p.y * (inv * p.zz);    // Y / ZZZ
    return a;
}

__device__ __forceinline__ Affine load_affine(const uint32_t *p)
{
    Affine a;
    a.x = load_fp(p);
    a.y = load_fp(p + 12);
    return a;
}

__device__ __forceinline__ void store_affine(uint32_t *p, const Affine &a)
{
    store_fp(p, a.x);
    store_fp(p + 12, a.y);
}

__device__ __forceinline__ Xyzz load_xyzz(const uint32_t *p)
{
    Xyzz a;
    a.x = load_fp(p);
    a.y = load_fp(p + 12);
    a.zz = load_fp(p + 24);
    a.zzz = load_fp(p + 36);
    return a;
}

__device__ __forceinline__ void store_xyzz(uint32_t *p, const Xyzz &a)
{
    store_fp(p, a.x);
    store_fp(p + 12, a.y);
    store_fp(p + 24, a.zz);
    store_fp(p + 36, a.zzz);
}

__device__ __forceinline__ void store_jacobian(uint32_t *p, const Jacobian &a)
{
    store_fp(p, a.x);
    store_fp(p + 12, a.y);
    store_fp(p + 24, a.z);
}

}  // namespace mh
