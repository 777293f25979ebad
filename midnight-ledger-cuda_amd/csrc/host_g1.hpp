// Host-side Fp and G1 Jacobian arithmetic for the short serial tails of the
// MSM (combining per-window bit-plane sums with a doubling chain), where one
// CPU core beats a single GPU lane by an order of magnitude.  Layout is the
// reference's RawG1Projective: x[6], y[6], z[6] u64 limbs, Montgomery form
// with R = 2^384, Z = 0 for the identity.
#pragma once
#include <cstdint>
#include <cstring>

namespace mh {

struct HFp {
    uint64_t l[6];
    static constexpr uint64_t P[6] = {0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL,
                                      0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};
    static constexpr uint64_t INV = 0x89f3fffcfffcfffdULL;   // -p^-1 mod 2^64
    static HFp zero() { return HFp{{0, 0, 0, 0, 0, 0}}; }
    static HFp one()
    {
        return HFp{{0x760900000002fffdULL, 0xebf4000bc40c0002ULL, 0x5f48985753c758baULL, 0x77ce585370525745ULL,
                    0x5c071a97a256ec6dULL, 0x15f65ec3fa80e493ULL}};
    }
    bool is_zero() const { return (l[0] | l[1] | l[2] | l[3] | l[4] | l[5]) == 0; }
    bool operator==(const HFp &o) const { return std::memcmp(l, o.l, 48) == 0; }

    static bool geq_p(const uint64_t *t)
    {
        for (int i = 5; i >= 0; --i) {
            if (t[i] > P[i]) return true;
            if (t[i] < P[i]) return false;
        }
        return true;
    }
    static void sub_p(uint64_t *t)
    {
        unsigned __int128 b = 0;
        for (int i = 0; i < 6; ++i) {
            unsigned __int128 d = (unsigned __int128)t[i] - P[i] - (uint64_t)b;
            t[i] = (uint64_t)d;
            b = (d >> 64) & 1;
        }
    }
    HFp operator+(const HFp &b) const
    {
        HFp r;
        unsigned __int128 c = 0;
        for (int i = 0; i < 6; ++i) {
            c = (unsigned __int128)l[i] + b.l[i] + (uint64_t)(c >> 64);
            r.l[i] = (uint64_t)c;
        }
        if (geq_p(r.l)) sub_p(r.l);   // p < 2^382: no carry out of the top limb
        return r;
    }
    HFp operator-(const HFp &b) const
    {
        HFp r;
        unsigned __int128 br = 0;
        for (int i = 0; i < 6; ++i) {
            unsigned __int128 d = (unsigned __int128)l[i] - b.l[i] - (uint64_t)br;
            r.l[i] = (uint64_t)d;
            br = (d >> 64) & 1;
        }
        if (br) {
            unsigned __int128 c = 0;
            for (int i = 0; i < 6; ++i) {
                c = (unsigned __int128)r.l[i] + P[i] + (uint64_t)(c >> 64);
                r.l[i] = (uint64_t)c;
            }
        }
        return r;
    }
    HFp operator*(const HFp &b) const
    {
        uint64_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 6; ++i) {
            unsigned __int128 c = 0;
            for (int j = 0; j < 6; ++j) {
                c = (unsigned __int128)l[j] * b.l[i] + t[j] + (uint64_t)(c >> 64);
                t[j] = (uint64_t)c;
            }
            unsigned __int128 s = (unsigned __int128)t[6] + (uint64_t)(c >> 64);
            t[6] = (uint64_t)s;
            t[7] = (uint64_t)(s >> 64);
            uint64_t m = t[0] * INV;
            c = (unsigned __int128)m * P[0] + t[0];
            for (int j = 1; j < 6; ++j) {
                c = (unsigned __int128)m * P[j] + t[j] + (uint64_t)(c >> 64);
                t[j - 1] = (uint64_t)c;
            }
            s = (unsigned __int128)t[6] + (uint64_t)(c >> 64);
            t[5] = (uint64_t)s;
            t[6] = t[7] + (uint64_t)(s >> 64);
        }
        if (t[6] || geq_p(t)) sub_p(t);
        HFp r;
        std::memcpy(r.l, t, 48);
        return r;
    }
    HFp sqr() const { return *this * *this; }
    HFp dbl() const { return *this + *this; }
    HFp inverse() const   // Fermat: a^(p-2); zero maps to zero
    {
        uint64_t e[6];
        std::memcpy(e, P, 48);
        e[0] -= 2;
        HFp acc = one(), base = *this;
        for (int i = 0; i < 384; ++i) {
            if ((e[i / 64] >> (i % 64)) & 1) acc = acc * base;
            base = base.sqr();
        }
        return acc;
    }
    void to_canonical(uint64_t out[6]) const
    {
        const HFp one_raw{{1, 0, 0, 0, 0, 0}};
        const HFp t = *this * one_raw;
        std::memcpy(out, t.l, 48);
    }
};

struct HJac {
    HFp x, y, z;
    static HJac identity() { return HJac{HFp::one(), HFp::one(), HFp::zero()}; }
    bool is_identity() const { return z.is_zero(); }

    // dbl-2009-l (a = 0)
    HJac dbl() const
    {
        if (is_identity() || y.is_zero()) return identity();
        HFp a = x.sqr(), b = y.sqr(), c = b.sqr();
        HFp d = ((x + b).sqr() - a - c).dbl();
        HFp e = a + a + a;
        HFp f = e.sqr();
        HJac r;
        r.x = f - d.dbl();
        HFp c8 = c.dbl().dbl().dbl();
        r.y = e * (d - r.x) - c8;
        r.z = (y * z).dbl();
        return r;
    }

    // add-2007-bl with the doubling / inverse cases
    HJac operator+(const HJac &q) const
    {
        if (is_identity()) return q;
        if (q.is_identity()) return *this;
        HFp z1z1 = z.sqr(), z2z2 = q.z.sqr();
        HFp u1 = x * z2z2, u2 = q.x * z1z1;
        HFp s1 = y * q.z * z2z2, s2 = q.y * z * z1z1;
        HFp h = u2 - u1, rr = (s2 - s1).dbl();
        if (h.is_zero()) {
            if (rr.is_zero()) return dbl();
            return identity();
        }
        HFp i = h.dbl().sqr();
        HFp j = h * i;
        HFp v = u1 * i;
        HJac r;
        r.x = rr.sqr() - j - v.dbl();
        r.y = rr * (v - r.x) - (s1 * j).dbl();
        r.z = ((z + q.z).sqr() - z1z1 - z2z2) * h;
        return r;
    }
};

}  // namespace mh
