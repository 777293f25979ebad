// omega^idx for any row index from a small table instead of a full
// powers vector: tab[l * 2048 + j] = omega^(j * 2^(11 l)) for j < 2048 and
// l < levels, levels = ceil(log2(len) / 11).  A row's power is the product
// of one entry per level (one product per row up to 2^22 rows), and the
// table (64 KB at 2^22) stays in cache.
#pragma once
#include <vector>

#include "fields.hpp"
#include "host_fr.hpp"
#include "runtime.hpp"

namespace mh {

constexpr int kOmegaBits = 11;
constexpr uint32_t kOmegaTab = 1u << kOmegaBits;

__device__ __forceinline__ Fr omega_pow_tab(size_t idx, const Fr *tab, uint32_t levels)
{
    Fr r = load_fr(tab + (idx & (kOmegaTab - 1)));
    for (uint32_t l = 1; l < levels; ++l)
        r = r * load_fr(tab + l * kOmegaTab + ((idx >> (kOmegaBits * l)) & (kOmegaTab - 1)));
    return r;
}

inline void build_omega_table(DevBuf<Fr> &tab, uint32_t &levels, const HostFr &omega, size_t len, hipStream_t st)
{
    uint32_t bits = 0;
    while (bits < 64 && (1ull << bits) < len) ++bits;
    levels = bits <= (uint32_t)kOmegaBits ? 1 : (bits + kOmegaBits - 1) / kOmegaBits;
    std::vector<HostFr> h((size_t)levels * kOmegaTab);
    HostFr step = omega;
    for (uint32_t l = 0; l < levels; ++l) {
        HostFr acc = HostFr::one();
        for (uint32_t j = 0; j < kOmegaTab; ++j) {
            h[(size_t)l * kOmegaTab + j] = acc;
            acc = acc * step;
        }
        step = acc;   // omega^(2048 * 2^(11 l))
    }
    tab.alloc(h.size());
    MH_HIP(hipMemcpyAsync(tab.get(), h.data(), h.size() * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipStreamSynchronize(st));
}

// The tables of npts points back to back (point p at p * levels *
// kOmegaTab), with one upload and one synchronisation for all of them.
inline void build_omega_tables(DevBuf<Fr> &tab, uint32_t &levels, const HostFr *points, size_t npts, size_t len,
                               hipStream_t st)
{
    uint32_t bits = 0;
    while (bits < 64 && (1ull << bits) < len) ++bits;
    levels = bits <= (uint32_t)kOmegaBits ? 1 : (bits + kOmegaBits - 1) / kOmegaBits;
    const size_t per = (size_t)levels * kOmegaTab;
    std::vector<HostFr> h(std::max<size_t>(npts * per, 1));
    for (size_t p = 0; p < npts; ++p) {
        HostFr step = points[p];
        for (uint32_t l = 0; l < levels; ++l) {
            HostFr acc = HostFr::one();
            for (uint32_t j = 0; j < kOmegaTab; ++j) {
                h[p * per + (size_t)l * kOmegaTab + j] = acc;
                acc = acc * step;
            }
            step = acc;
        }
    }
    tab.alloc(h.size());
    MH_HIP(hipMemcpyAsync(tab.get(), h.data(), h.size() * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipStreamSynchronize(st));
}

}  // namespace mh
