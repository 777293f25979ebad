// omega^idx for any row index from a small table instead of a full
// powers vector: tab[l * 2048 + j] = omega^(j * 2^(11 l)) for j < 2048 and
// l < levels, levels = ceil(log2(len) / 11).  A row's power is the product
// of one entry per level (one product per row up to 2^22 rows), and the
// table (64 KB at 2^22) stays in cache.
#pragma once
#include <algorithm>
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

// tab[(p * levels + l) * kOmegaTab + j] = points[p]^(j * 2^(11 l)), computed
// on the device (square-and-multiply per entry, no upload and no wait)
void omega_tables_device(Fr *tab, uint32_t levels, const HostFr *points, size_t npts, hipStream_t st);

inline uint32_t omega_levels(size_t len)
{
    uint32_t bits = 0;
    while (bits < 64 && (1ull << bits) < len) ++bits;
    return bits <= (uint32_t)kOmegaBits ? 1 : (bits + kOmegaBits - 1) / kOmegaBits;
}

inline void build_omega_table(DevBuf<Fr> &tab, uint32_t &levels, const HostFr &omega, size_t len, hipStream_t st)
{
    levels = omega_levels(len);
    tab.alloc((size_t)levels * kOmegaTab);
    omega_tables_device(tab.get(), levels, &omega, 1, st);
}

// The tables of npts points back to back (point p at p * levels *
// kOmegaTab).
inline void build_omega_tables(DevBuf<Fr> &tab, uint32_t &levels, const HostFr *points, size_t npts, size_t len,
                               hipStream_t st)
{
    levels = omega_levels(len);
    tab.alloc(std::max<size_t>(npts * levels * kOmegaTab, 1));
    omega_tables_device(tab.get(), levels, points, npts, st);
}

}  // namespace mh
