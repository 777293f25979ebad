// Host-side Fr arithmetic for the few constants the backend derives itself
// (roots of unity, twiddle tables, 1/n).  Same representation as RawFr:
// 4 x u64 little-endian limbs, Montgomery form with R = 2^256.
#pragma once
#include <cstdint>
#include <cstring>

namespace mh {

struct HostFr {
    uint64_t l[4];

    static constexpr uint64_t P[4] = {0xffffffff00000001ULL, 0x53bda402fffe5bfeULL,
                                      0x3339d80809a1d805ULL, 0x73eda753299d7d48ULL};
    static constexpr uint64_t INV = 0xfffffffeffffffffULL;   // -r^-1 mod 2^64
    static constexpr uint64_t R2[4] = {0xc999e990f3f29c6dULL, 0x2b6cedcb87925c23ULL,
                                       0x05d314967254398fULL, 0x0748d9d99f59ff11ULL};

    static HostFr zero() { return HostFr{{0, 0, 0, 0}}; }
    static HostFr one()
    {
        return HostFr{{0x00000001fffffffeULL, 0x5884b7fa00034802ULL, 0x998c4fefecbc4ff5ULL,
                       0x1824b159acc5056fULL}};
    }
    // 2^32-th root of unity (midnight-curves fq.rs:77-84), Montgomery form
    static HostFr root_of_unity()
    {
        return HostFr{{0xb9b58d8c5f0e466aULL, 0x5b1b4c801819d7ecULL, 0x0af53ae352a31e64ULL,
                       0x5bf3adda19e9b27bULL}};
    }

    bool operator==(const HostFr &o) const { return std::memcmp(l, o.l, 32) == 0; }

    static bool geq_p(const uint64_t *t)
    {
        for (int i = 3; i >= 0; --i) {
            if (t[i] > P[i]) return true;
            if (t[i] < P[i]) return false;
        }
        return true;
    }
    static void sub_p(uint64_t *t)
    {
        unsigned __int128 b = 0;
        for (int i = 0; i < 4; ++i) {
            unsigned __int128 d = (unsigned __int128)t[i] - P[i] - (uint64_t)b;
            t[i] = (uint64_t)d;
            b = (d >> 64) & 1;
        }
    }

    HostFr operator*(const HostFr &b) const
    {
        uint64_t t[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 4; ++i) {
            unsigned __int128 c = 0;
            for (int j = 0; j < 4; ++j) {
                c = (unsigned __int128)l[j] * b.l[i] + t[j] + (uint64_t)(c >> 64);
                t[j] = (uint64_t)c;
            }
            unsigned __int128 s = (unsigned __int128)t[4] + (uint64_t)(c >> 64);
            t[4] = (uint64_t)s;
            t[5] = (uint64_t)(s >> 64);
            uint64_t m = t[0] * INV;
            c = (unsigned __int128)m * P[0] + t[0];
            for (int j = 1; j < 4; ++j) {
                c = (unsigned __int128)m * P[j] + t[j] + (uint64_t)(c >> 64);
                t[j - 1] = (uint64_t)c;
            }
            s = (unsigned __int128)t[4] + (uint64_t)(c >> 64);
            t[3] = (uint64_t)s;
            t[4] = t[5] + (uint64_t)(s >> 64);
        }
        if (t[4] || geq_p(t)) sub_p(t);
        HostFr r;
        std::memcpy(r.l, t, 32);
        return r;
    }

    HostFr sqr() const { return *this * *this; }

    HostFr operator+(const HostFr &b) const
    {
        HostFr r;
        unsigned __int128 c = 0;
        for (int i = 0; i < 4; ++i) {
            c = (unsigned __int128)l[i] + b.l[i] + (uint64_t)(c >> 64);
            r.l[i] = (uint64_t)c;
        }
        if (geq_p(r.l)) sub_p(r.l);   // r < 2^255: no carry out of the top limb
        return r;
    }
    HostFr operator-(const HostFr &b) const
    {
        HostFr r;
        unsigned __int128 br = 0;
        for (int i = 0; i < 4; ++i) {
            unsigned __int128 d = (unsigned __int128)l[i] - b.l[i] - (uint64_t)br;
            r.l[i] = (uint64_t)d;
            br = (d >> 64) & 1;
        }
        if (br) {
            unsigned __int128 c = 0;
            for (int i = 0; i < 4; ++i) {
                c = (unsigned __int128)r.l[i] + P[i] + (uint64_t)(c >> 64);
                r.l[i] = (uint64_t)c;
            }
        }
        return r;
    }
    HostFr neg() const { return zero() - *this; }
    bool is_zero() const { return (l[0] | l[1] | l[2] | l[3]) == 0; }

    // canonical integer < r  ->  Montgomery form
    static HostFr from_canonical(const uint64_t c[4])
    {
        HostFr t{{c[0], c[1], c[2], c[3]}};
        HostFr r2;
        std::memcpy(r2.l, R2, 32);
        return t * r2;
    }
    void to_canonical(uint64_t out[4]) const
    {
        const HostFr one_raw{{1, 0, 0, 0}};
        const HostFr t = *this * one_raw;
        std::memcpy(out, t.l, 32);
    }

    HostFr pow(const uint64_t e[4]) const
    {
        HostFr acc = one(), base = *this;
        for (int i = 0; i < 256; ++i) {
            if ((e[i / 64] >> (i % 64)) & 1) acc = acc * base;
            base = base.sqr();
        }
        return acc;
    }

    HostFr inverse() const
    {
        const uint64_t e[4] = {P[0] - 2, P[1], P[2], P[3]};   // Fermat: a^(r-2)
        return pow(e);
    }

    static HostFr from_u64(uint64_t v)
    {
        HostFr t{{v, 0, 0, 0}};
        HostFr r2;
        std::memcpy(r2.l, R2, 32);
        return t * r2;
    }

    // primitive 2^log_n-th root of unity used by the prover's domains
    // (domain.rs:71-86): ROOT_OF_UNITY^(2^(32 - log_n))
    static HostFr omega(uint32_t log_n)
    {
        HostFr w = root_of_unity();
        for (uint32_t i = log_n; i < 32; ++i) w = w.sqr();
        return w;
    }
};

}  // namespace mh
