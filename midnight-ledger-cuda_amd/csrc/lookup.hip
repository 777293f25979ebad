// Lookup-argument permutation on the device: the permuted input and table
// columns of halo2's permute_expression_pair
// (vendor/midnight-proofs-0.8.1/src/plonk/lookup/prover.rs:1573-1690),
// behind fr_sort_canonical (bls12_381_g1_msm.cu:6617),
// fr_lookup_permute_batch (:5139) and trace_context_lookup_permute (:5849).
//
// Per lookup, over the usable rows:
//   permuted input = the inputs sorted by canonical value (Fq's Ord);
//   each first occurrence of a value takes the same value in the permuted
//   table, and removes the first equal entry of the sorted table (its lower
//   bound: the unique inputs are ascending, so this is the entry the CPU's
//   merge pointer reaches); an input value absent from the table is an
//   error;
//   the table entries left over, ascending, fill the repeated rows from the
//   last one back.
// Canonical order is a 256-bit key: four stable LSD passes of a 64-bit
// radix sort carrying the element index.  Compactions (leftovers, repeated
// rows) are exclusive scans of flags.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "devvec.hpp"
#include "fields.hpp"
#include "lookup.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_of(size_t n) { return (unsigned)std::max<size_t>(ceil_div(n, kBlock), 1); }

struct Canon {
    uint64_t l[4];
};

__global__ void k_canon(Canon *keys, const Fr *v, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Fr c = fr_from_mont(load_fr(v + i));
    Canon k;
#pragma unroll
    for (int j = 0; j < 4; ++j) k.l[j] = (uint64_t)c.v[2 * j] | (uint64_t)c.v[2 * j + 1] << 32;
    keys[i] = k;
}

__global__ void k_iota(uint32_t *idx, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) idx[i] = (uint32_t)i;
}

__global__ void k_limb(uint64_t *out, const Canon *keys, const uint32_t *idx, size_t n, int limb)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = keys[idx[i]].l[limb];
}

__global__ void k_gather_sorted(Fr *vals_out, Canon *keys_out, const Fr *vals, const Canon *keys,
                                const uint32_t *idx, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    store_fr(vals_out + i, load_fr(vals + idx[i]));
    if (keys_out) keys_out[i] = keys[idx[i]];
}

__device__ __forceinline__ int canon_cmp(const Canon &a, const Canon &b)
{
    for (int j = 3; j >= 0; --j) {
        if (a.l[j] < b.l[j]) return -1;
        if (a.l[j] > b.l[j]) return 1;
    }
    return 0;
}

// row r of the sorted inputs: first occurrence (1) or repeat; first
// occurrences give the permuted table its value at r
__global__ void k_mark_inputs(const Canon *keys, const Fr *sorted, uint32_t *uniq, uint32_t *dup, Fr *ptable,
                              size_t n)
{
    const size_t r = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    const bool u = r == 0 || canon_cmp(keys[r], keys[r - 1]) != 0;
    uniq[r] = u;
    dup[r] = !u;
    if (u) store_fr(ptable + r, load_fr(sorted + r));
}

// each first occurrence removes the lower bound of its value in the sorted
// table (leftover flag 0); absent values set the error flag
__global__ void k_match(const Canon *in_keys, const uint32_t *uniq, const Canon *tab_keys, uint32_t *leftover,
                        uint32_t *invalid, size_t n)
{
    const size_t r = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (r >= n || !uniq[r]) return;
    const Canon v = in_keys[r];
    size_t lo = 0, hi = n;
    while (lo < hi) {
        const size_t mid = lo + (hi - lo) / 2;
        if (canon_cmp(tab_keys[mid], v) < 0) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= n || canon_cmp(tab_keys[lo], v) != 0) {
        atomicOr(invalid, 1u);
        return;
    }
    leftover[lo] = 0;
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// repeated rows, ascending: rows[dpos[r]] = r
__global__ void k_dup_rows(const uint32_t *dup, const uint32_t *dpos, uint32_t *rows, size_t n)
{
    const size_t r = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (r < n && dup[r]) rows[dpos[r]] = (uint32_t)r;
}

// leftover k (ascending) -> repeated row rows[D - 1 - k]; the counts must agree
__global__ void k_assign_leftovers(const uint32_t *leftover, const uint32_t *lpos, const uint32_t *dup,
                                   const uint32_t *dpos, const uint32_t *rows, const Fr *sorted_table, Fr *ptable,
                                   uint32_t *invalid, size_t n)
{
    const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t D = dpos[n - 1] + dup[n - 1];
    const uint32_t Lc = lpos[n - 1] + leftover[n - 1];
    if (D != Lc) {
        if (p == 0) atomicOr(invalid, 1u);
        return;
    }
    if (!leftover[p]) return;
    const uint32_t k = lpos[p];
    store_fr(ptable + rows[D - 1 - k], load_fr(sorted_table + p));
}

// Working buffers of one lane, grown as needed.
struct LookupScratch {
    DevBuf<Canon> keys, keys2, in_keys;
    DevBuf<uint64_t> limb, limb_out;
    DevBuf<uint32_t> idx, idx_out, f1, f2, p1, p2, rows, invalid;
    DevBuf<Fr> sorted;
    DevBuf<uint8_t> temp;
};

LookupScratch &lookup_scratch()
{
    static LookupScratch s[64][kLaneCount];
    return s[device_id()][current_lane()];
}

void ensure_temp(LookupScratch &sc, size_t n, hipStream_t st)
{
    size_t a = 0, b = 0;
    MH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, a, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                              (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 64, st));
    MH_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, st));
    sc.temp.ensure(std::max<size_t>(std::max(a, b), 1));
}

// sorted values (and their canonical keys) of n values
void sort_canonical(Fr *out, Canon *out_keys, const Fr *v, size_t n, LookupScratch &sc, hipStream_t st)
{
    sc.keys.ensure(n);
    sc.limb.ensure(n);
    sc.limb_out.ensure(n);
    sc.idx.ensure(n);
    sc.idx_out.ensure(n);
    ensure_temp(sc, n, st);
    hipLaunchKernelGGL(k_canon, dim3(grid_of(n)), dim3(kBlock), 0, st, sc.keys.get(), v, n);
    hipLaunchKernelGGL(k_iota, dim3(grid_of(n)), dim3(kBlock), 0, st, sc.idx.get(), n);
    MH_HIP(hipGetLastError());
    for (int limb = 0; limb < 4; ++limb) {   // least significant limb first; the sort is stable
        hipLaunchKernelGGL(k_limb, dim3(grid_of(n)), dim3(kBlock), 0, st, sc.limb.get(), (const Canon *)sc.keys.get(),
                           (const uint32_t *)sc.idx.get(), n, limb);
        MH_HIP(hipGetLastError());
        size_t bytes = sc.temp.len;
        MH_HIP(hipcub::DeviceRadixSort::SortPairs(sc.temp.get(), bytes, (const uint64_t *)sc.limb.get(),
                                                  sc.limb_out.get(), (const uint32_t *)sc.idx.get(),
                                                  sc.idx_out.get(), (int)n, 0, 64, st));
        std::swap(sc.idx, sc.idx_out);
    }
    hipLaunchKernelGGL(k_gather_sorted, dim3(grid_of(n)), dim3(kBlock), 0, st, out, out_keys, v,
                       (const Canon *)sc.keys.get(), (const uint32_t *)sc.idx.get(), n);
    MH_HIP(hipGetLastError());
}

void exclusive_sum(uint32_t *out, const uint32_t *in, size_t n, LookupScratch &sc, hipStream_t st)
{
    size_t bytes = sc.temp.len;
    MH_HIP(hipcub::DeviceScan::ExclusiveSum(sc.temp.get(), bytes, in, out, (int)n, st));
}

}  // namespace

void fr_sort_canonical_device(Fr *out, const Fr *v, size_t n, hipStream_t st)
{
    if (n == 0) return;
    MH_CHECK(n < (1ull << 31), (int)hipErrorInvalidValue, "too many values to sort");
    LookupScratch &sc = lookup_scratch();
    sort_canonical(out, nullptr, v, n, sc, st);
    MH_HIP(hipStreamSynchronize(st));
}

bool lookup_permute_device(Fr *trace, size_t len, size_t lookup_count, size_t usable, const Fr *input_blinds,
                           const Fr *table_blinds, size_t blind_count, hipStream_t st)
{
    MH_CHECK(usable > 0 && usable <= len && usable < (1ull << 31) && blind_count == len - usable,
             (int)hipErrorInvalidValue, "invalid lookup permutation shape");
    const size_t group = lookup_count * len;
    LookupScratch &sc = lookup_scratch();
    const size_t n = usable;
    sc.in_keys.ensure(n);
    sc.keys2.ensure(n);
    sc.sorted.ensure(n);
    sc.f1.ensure(n);
    sc.f2.ensure(n);
    sc.p1.ensure(n);
    sc.p2.ensure(n);
    sc.rows.ensure(n);
    sc.invalid.ensure(1);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, sc.invalid.get(), (size_t)1, 0u);
    MH_HIP(hipGetLastError());
    const dim3 g(grid_of(n)), b(kBlock);
    for (size_t l = 0; l < lookup_count; ++l) {
        const Fr *input = trace + l * len;
        const Fr *table = trace + group + l * len;
        Fr *pin = trace + 2 * group + l * len;
        Fr *ptab = trace + 3 * group + l * len;
        // permuted input: the sorted inputs; first occurrences to the table side
        sort_canonical(pin, sc.in_keys.get(), input, n, sc, st);
        hipLaunchKernelGGL(k_mark_inputs, g, b, 0, st, (const Canon *)sc.in_keys.get(), (const Fr *)pin,
                           sc.f1.get(), sc.f2.get(), ptab, n);   // f1 unique, f2 repeated
        MH_HIP(hipGetLastError());
        // sorted table; leftover flags start at 1 and matches clear them
        sort_canonical(sc.sorted.get(), sc.keys2.get(), table, n, sc, st);
        hipLaunchKernelGGL(k_fill_u32, g, b, 0, st, sc.p2.get(), n, 1u);   // p2: leftover flags
        hipLaunchKernelGGL(k_match, g, b, 0, st, (const Canon *)sc.in_keys.get(), (const uint32_t *)sc.f1.get(),
                           (const Canon *)sc.keys2.get(), sc.p2.get(), sc.invalid.get(), n);
        MH_HIP(hipGetLastError());
        // positions: repeated rows (f2 -> p1) and leftovers (p2 -> f1, the
        // unique flags are no longer needed)
        exclusive_sum(sc.p1.get(), sc.f2.get(), n, sc, st);
        exclusive_sum(sc.f1.get(), sc.p2.get(), n, sc, st);
        hipLaunchKernelGGL(k_dup_rows, g, b, 0, st, (const uint32_t *)sc.f2.get(), (const uint32_t *)sc.p1.get(),
                           sc.rows.get(), n);
        hipLaunchKernelGGL(k_assign_leftovers, g, b, 0, st, (const uint32_t *)sc.p2.get(),
                           (const uint32_t *)sc.f1.get(), (const uint32_t *)sc.f2.get(),
                           (const uint32_t *)sc.p1.get(), (const uint32_t *)sc.rows.get(),
                           (const Fr *)sc.sorted.get(), ptab, sc.invalid.get(), n);
        MH_HIP(hipGetLastError());
        if (blind_count) {
            MH_HIP(hipMemcpyAsync(pin + usable, input_blinds + l * blind_count, blind_count * 32,
                                  hipMemcpyDeviceToDevice, st));
            MH_HIP(hipMemcpyAsync(ptab + usable, table_blinds + l * blind_count, blind_count * 32,
                                  hipMemcpyDeviceToDevice, st));
        }
    }
    uint32_t bad = 0;
    MH_HIP(hipMemcpyAsync(&bad, sc.invalid.get(), 4, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return bad == 0;
}

}  // namespace mh

// ---------------------------------------------------------------- C ABI

using namespace mh;

// bls12_381_g1_msm.cu:6617-6676: values sorted by canonical value, in place
extern "C" RustError midnight_sppark_bls12_381_fr_sort_canonical(RawFr *values, size_t len)
{
    if (len < 2) return ok();
    if (!values) return make_error((int)hipErrorInvalidValue, "values is null");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> in(len), out(len);
    MH_HIP(hipMemcpyAsync(in.get(), values, len * 32, hipMemcpyHostToDevice, st));
    fr_sort_canonical_device(out.get(), in.get(), len, st);
    MH_HIP(hipMemcpyAsync(values, out.get(), len * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:5139-5194: a device vector [compressed inputs |
// compressed tables | permuted inputs | permuted tables] from host
// compressed columns (lookup-major, len rows each)
extern "C" RustError midnight_sppark_bls12_381_fr_lookup_permute_batch(
    void **out, const RawFr *compressed_inputs, const RawFr *compressed_tables, size_t len, size_t lookup_count,
    size_t usable_rows, const RawFr *input_blinds, const RawFr *table_blinds, size_t blind_count)
{
    if (out == nullptr || compressed_inputs == nullptr || compressed_tables == nullptr || input_blinds == nullptr ||
        table_blinds == nullptr || len == 0 || lookup_count == 0 || usable_rows == 0 || usable_rows > len ||
        blind_count != len - usable_rows || usable_rows > UINT32_MAX || lookup_count > SIZE_MAX / len)
        return make_error((int)hipErrorInvalidValue, "invalid resident lookup permutation arguments");
    *out = nullptr;
    const size_t group = lookup_count * len;
    if (group > SIZE_MAX / 4) return make_error((int)hipErrorInvalidValue, "resident lookup permutation size overflow");
    MH_TRY
    hipStream_t st = cur_stream();
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = 4 * group;
    dv->values.alloc(dv->len);
    Fr *t = dv->values.get();
    MH_HIP(hipMemcpyAsync(t, compressed_inputs, group * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(t + group, compressed_tables, group * 32, hipMemcpyHostToDevice, st));
    DevBuf<Fr> blinds(std::max<size_t>(2 * lookup_count * blind_count, 1));
    if (blind_count) {
        MH_HIP(hipMemcpyAsync(blinds.get(), input_blinds, lookup_count * blind_count * 32, hipMemcpyHostToDevice, st));
        MH_HIP(hipMemcpyAsync(blinds.get() + lookup_count * blind_count, table_blinds, lookup_count * blind_count * 32,
                              hipMemcpyHostToDevice, st));
    }
    if (!lookup_permute_device(t, len, lookup_count, usable_rows, blinds.get(),
                               blinds.get() + lookup_count * blind_count, blind_count, st))
        return make_error(-1, "lookup input is not contained in the lookup table");
    *out = dv.release();
    return ok();
    MH_CATCH
}
