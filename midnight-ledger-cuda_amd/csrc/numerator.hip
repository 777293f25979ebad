// Quotient numerator on the extended domain.
//
// One fused pass per evaluation: every lane owns one row of the evaluated
// domain and folds, in the reference's order, the custom-gate graph, the
// permutation terms, each lookup's five terms and each trash term into one
// value (value = value * y + term), which it writes once.  The reference
// runs one kernel per phase and re-reads / re-writes the output between
// them (bls12_381_g1_msm.cu:620-875 numerator_{custom,permutation,lookup,
// trash}_kernel, launched by :5196 eval_numerator and :5968
// numerator_context_eval); the values produced are the same.
//
// The expression graphs are interpreted: the program (constants, rotations,
// calculations, Horner parts) is identical for every row, so its fields are
// read with wave-uniform loads and every branch on an opcode or source kind
// is uniform.  Each lane keeps the graph's intermediate slots in LDS,
// laid out [slot][word][lane] so that the 64 lanes of a wave touch 64
// consecutive words (no bank conflicts).  Graphs with more slots than fit
// there keep them in a global slot file of the same layout.
//
// omega^idx for the permutation identity columns comes from a two- or
// three-level table (omega^(j * 2^(11 l)), j < 2048) instead of a full
// prefix-product vector: one or two products per row and no per-row HBM
// read.
//
// HBM traffic per row for the synthetic circuit (5 advice, 3 fixed, 8
// sigma, 3 permutation product columns, l0 / l_last / l_active, one output):
// 23 x 32 B = 736 B compulsory; rotated reads of the product columns
// (r_next, r_last) fall on lines that neighbouring rows read too.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "devvec.hpp"
#include "fields.hpp"
#include "fr29.hpp"
#include "grandproduct.hpp"
#include "lookup.hpp"
#include "numerator.hpp"
#include "host_fr.hpp"
#include "profile.hpp"
#include "rowpow.hpp"
#include "runtime.hpp"

namespace mh {

// Global slot file for graphs with more intermediate slots than LDS holds;
// one per lane and device, grown on demand and reused.
struct LaneSlotFile {
    DevBuf<uint32_t> buf;
};
LaneSlotFile &numerator_slot_file();

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMaxSlots = 256;       // MAX_GRAPH_SCRATCH_SLOTS (midnight-gpu-prover expression.rs:23)
constexpr uint32_t kLdsSlots = 4;         // 4 x 36 B x 256 lanes = 36 KB of LDS per block
constexpr uint32_t kGlobalSlotBlocks = 1024;

struct DevGraph {
    const Fr *constants;   // scaled form (c 2^261 mod r), converted when the graph is uploaded
    const int32_t *rotations;
    const RawCalculation *calcs;
    const RawValueSource *parts;
    uint32_t constants_len, rotations_len, calcs_len, parts_len, num_slots;
};

// Which inputs of an evaluation are already stored in the scaled form
// (v 2^261 mod r); the others are in the usual form and converted as they
// are loaded.
enum : uint32_t {
    kScFixed = 1, kScAdvice = 2, kScInstance = 4, kScProducts = 8, kScCosets = 16, kScLagrange = 32,
    kScLookup = 64, kScTrash = 128, kScInitial = 256
};

struct NumArgs {
    Fr *out;                  // usual form
    const Fr *initial;
    size_t len;
    const Fr *fixed, *advice, *instance, *challenges;   // challenges: scaled
    uint32_t fixed_cols, advice_cols, instance_cols, challenge_len;
    Fr beta, gamma, theta, trash, y, delta_start, delta_const;   // scaled
    int32_t rot_scale, last_rotation;
    uint32_t chunk_len, run_custom;
    const Fr *l0, *l_last, *l_active;
    const Fr *perm_products;
    const RawColumnRef *perm_columns;
    const Fr *perm_cosets;
    uint32_t perm_sets, perm_cols;
    const Fr *omega_tab;      // scaled; level 0 times delta_start when omega_has_delta
    uint32_t omega_levels, omega_has_delta;
    uint32_t out_scaled;      // write the scaled form (fully reduced) instead of the usual one
    uint32_t lookup_count, trash_count;
    const Fr *lookup_products, *lookup_inputs, *lookup_tables;
    const Fr *trash_cosets;
    const uint32_t *trash_sel;
    const DevGraph *graphs;   // [custom, lookup 0.., trash 0..]
    uint32_t *slots_global;   // [slot * 9 + limb][slot_stride] when not in LDS
    size_t slot_stride;
    uint32_t graph_count;     // graphs evaluated by k_eval_graphs
    uint32_t scaled;          // kSc* flags
};

// ---------------------------------------------------------------- device
//
// Every value inside the kernels is in the scaled form of fr29.hpp: nine
// 29-bit limbs holding v 2^261 mod r, loose (any representative below 2r),
// so that each product is one radix-2^29 Montgomery product (v_mad_u64_u32
// on 64-bit accumulators, no carry chains, no final subtraction) and stays
// in that form.  Results leave fully reduced in the usual form, so they are
// limb-identical to the reference's.

using S = Fr29;

__device__ __forceinline__ S s_zero()
{
    S r;
#pragma unroll
    for (int j = 0; j < 9; ++j) r.v[j] = 0;
    return r;
}
__device__ __forceinline__ S s_add(const S &a, const S &b) { return fr29_add_loose(a, b); }
__device__ __forceinline__ S s_sub(const S &a, const S &b) { return fr29_sub_loose(a, b); }
__device__ __forceinline__ S s_mul(const S &a, const S &b) { return fr29_mul_loose(a, b); }
// v y + t l, the numerator's fold step, with one reduction for both products
__device__ __forceinline__ S s_fold(const S &v, const S &y, const S &t, const S &l)
{
    return fr29_mul2_loose(v, y, t, l);
}
// v y + (p - q) l for loose p, q: the difference stays lazy (p - q + 4r,
// below 6r, limbs below 2^31), so each of the product's rows adds at most
// 2^58 + 2^60 + 2^58 and its accumulators stay below 54 x 2^58 < 2^64; the
// result is below (2r r + 6r r) / 2^261 + r < 1.12r, loose
__device__ __forceinline__ S s_fold_diff(const S &v, const S &y, const S &p, const S &q, const S &l)
{
    return fr29_mul2_loose(v, y, fr29_lazy_sub4r(p, q), l);
}
__device__ __forceinline__ S s_const(const Fr &c) { return fr29_from32(c); }   // already scaled
__device__ __forceinline__ S s_load(const Fr *p, bool scaled)
{
    const Fr x = load_fr(p);
    return scaled ? fr29_from32(x) : s29_from_usual(x);
}

__device__ __forceinline__ size_t rotated(size_t idx, int32_t rot, int32_t scale, size_t len)
{
    int64_t s = (int64_t)idx + (int64_t)rot * (int64_t)scale;
    if (s < 0 || s >= (int64_t)len) {
        s %= (int64_t)len;
        if (s < 0) s += (int64_t)len;
    }
    return (size_t)s;
}

// column `index` of kind (0 fixed, 1 advice, 2 instance), or nullptr; sets
// whether that kind is stored scaled
__device__ __forceinline__ const Fr *column_base(uint32_t kind, uint32_t index, const NumArgs &a, bool &scaled)
{
    if (kind == 0 && index < a.fixed_cols) {
        scaled = a.scaled & kScFixed;
        return a.fixed + (size_t)index * a.len;
    }
    if (kind == 1 && index < a.advice_cols) {
        scaled = a.scaled & kScAdvice;
        return a.advice + (size_t)index * a.len;
    }
    if (kind == 2 && index < a.instance_cols) {
        scaled = a.scaled & kScInstance;
        return a.instance + (size_t)index * a.len;
    }
    return nullptr;
}

// Per-lane slot file: limb w of slot s at [(s * 9 + w) * stride + lane].
// Slots<true> addresses the block's dynamic LDS (stride = kBlock lanes),
// Slots<false> the global slot file (stride = lanes of the whole grid).
extern __shared__ uint32_t lds_slots[];

template <bool kLds>
struct Slots;

template <>
struct Slots<true> {
    uint32_t lane;
    __device__ __forceinline__ S load(uint32_t s) const
    {
        S r;
#pragma unroll
        for (int w = 0; w < 9; ++w) r.v[w] = lds_slots[(s * 9 + w) * kBlock + lane];
        return r;
    }
    __device__ __forceinline__ void store(uint32_t s, const S &v) const
    {
#pragma unroll
        for (int w = 0; w < 9; ++w) lds_slots[(s * 9 + w) * kBlock + lane] = v.v[w];
    }
};

template <>
struct Slots<false> {
    uint32_t *base;
    size_t stride;
    size_t lane;
    __device__ __forceinline__ S load(uint32_t s) const
    {
        S r;
#pragma unroll
        for (int w = 0; w < 9; ++w) r.v[w] = base[((size_t)s * 9 + w) * stride + lane];
        return r;
    }
    __device__ __forceinline__ void store(uint32_t s, const S &v) const
    {
#pragma unroll
        for (int w = 0; w < 9; ++w) base[((size_t)s * 9 + w) * stride + lane] = v.v[w];
    }
};

template <bool kLds>
__device__ __forceinline__ S source_value(const RawValueSource &s, const DevGraph &g, size_t idx, const NumArgs &a,
                                          const Slots<kLds> &sl)
{
    switch (s.kind) {
    case 0: return s.a < g.constants_len ? s_const(load_fr(g.constants + s.a)) : s_zero();
    case 1: return s.a < g.num_slots ? sl.load(s.a) : s_zero();
    case 2:
    case 3:
    case 4: {
        bool scaled = false;
        const Fr *base = column_base(s.kind - 2, s.a, a, scaled);
        if (!base || s.b >= g.rotations_len) return s_zero();
        return s_load(base + rotated(idx, g.rotations[s.b], a.rot_scale, a.len), scaled);
    }
    case 5: return s.a < a.challenge_len ? s_const(load_fr(a.challenges + s.a)) : s_zero();
    case 6: return s_const(a.beta);
    case 7: return s_const(a.gamma);
    case 8: return s_const(a.theta);
    case 9: return s_const(a.trash);
    case 10: return s_const(a.y);
    default: return s_zero();   // 11: previous value, zero in every evaluation
    }
}

template <bool kLds>
__device__ __forceinline__ S eval_graph(const DevGraph &g, size_t idx, const NumArgs &a, const Slots<kLds> &sl)
{
    for (uint32_t s = 0; s < g.num_slots; ++s) sl.store(s, s_zero());
    S last = s_zero();
    for (uint32_t c = 0; c < g.calcs_len; ++c) {
        const RawCalculation calc = g.calcs[c];
        S v;
        switch (calc.op) {
        case 0: v = s_add(source_value(calc.lhs, g, idx, a, sl), source_value(calc.rhs, g, idx, a, sl)); break;
        case 1: v = s_sub(source_value(calc.lhs, g, idx, a, sl), source_value(calc.rhs, g, idx, a, sl)); break;
        case 2: v = s_mul(source_value(calc.lhs, g, idx, a, sl), source_value(calc.rhs, g, idx, a, sl)); break;
        case 3: {
            const S x = source_value(calc.lhs, g, idx, a, sl);
            v = s_mul(x, x);
            break;
        }
        case 4: {
            const S x = source_value(calc.lhs, g, idx, a, sl);
            v = s_add(x, x);
            break;
        }
        case 5: v = s_sub(s_zero(), source_value(calc.lhs, g, idx, a, sl)); break;
        case 6: {
            const S f = source_value(calc.rhs, g, idx, a, sl);
            v = source_value(calc.lhs, g, idx, a, sl);
            for (uint32_t p = 0; p < calc.parts_len; ++p) {
                const uint32_t pi = calc.parts_offset + p;
                const S part = pi < g.parts_len ? source_value(g.parts[pi], g, idx, a, sl) : s_zero();
                v = s_add(s_mul(v, f), part);
            }
            break;
        }
        case 7: v = source_value(calc.lhs, g, idx, a, sl); break;
        default: v = s_zero(); break;
        }
        if (calc.target < g.num_slots) sl.store(calc.target, v);
        last = v;
    }
    return last;
}

// omega^idx from the scaled two- or three-level table (rowpow.hpp layout)
__device__ __forceinline__ S omega_pow(size_t idx, const NumArgs &a)
{
    S r = s_const(load_fr(a.omega_tab + (idx & (kOmegaTab - 1))));
    for (uint32_t l = 1; l < a.omega_levels; ++l)
        r = s_mul(r, s_const(load_fr(a.omega_tab + l * kOmegaTab + ((idx >> (kOmegaBits * l)) & (kOmegaTab - 1)))));
    return r;
}

// Permutation terms (numerator_permutation_kernel, bls12_381_g1_msm.cu:674-755;
// evaluation.rs:3283-3352).
__device__ __forceinline__ S permutation_terms(S v, size_t idx, const NumArgs &a)
{
    const size_t len = a.len;
    const size_t r_next = rotated(idx, 1, a.rot_scale, len);
    const size_t r_last = rotated(idx, a.last_rotation, a.rot_scale, len);
    const bool zs = a.scaled & kScProducts, ls = a.scaled & kScLagrange, cs = a.scaled & kScCosets;
    const S one = s29_one();
    const S y = s_const(a.y);
    const S l0 = s_load(a.l0 + idx, ls);
    const Fr *first = a.perm_products;
    const Fr *last = a.perm_products + (size_t)(a.perm_sets - 1) * len;

    v = s_fold_diff(v, y, one, s_load(first + idx, zs), l0);
    const S zl = s_load(last + idx, zs);
    v = s_fold_diff(v, y, s_mul(zl, zl), zl, s_load(a.l_last + idx, ls));
    for (uint32_t s = 1; s < a.perm_sets; ++s) {
        const S cur = s_load(a.perm_products + (size_t)s * len + idx, zs);
        const S prev = s_load(a.perm_products + (size_t)(s - 1) * len + r_last, zs);
        v = s_fold_diff(v, y, cur, prev, l0);
    }

    const S l_active = s_load(a.l_active + idx, ls);
    const S beta = s_const(a.beta), gamma = s_const(a.gamma), dconst = s_const(a.delta_const);
    // delta_start omega^idx (the table's first level may already hold delta_start)
    S delta = a.omega_has_delta ? omega_pow(idx, a) : s_mul(s_const(a.delta_start), omega_pow(idx, a));
    for (uint32_t s = 0; s < a.perm_sets; ++s) {
        const Fr *z = a.perm_products + (size_t)s * len;
        S left = s_load(z + r_next, zs);
        S right = s_load(z + idx, zs);
        const uint32_t base = s * a.chunk_len;
        for (uint32_t j = 0; j < a.chunk_len && base + j < a.perm_cols; ++j) {
            const RawColumnRef c = a.perm_columns[base + j];
            bool scaled = false;
            const Fr *col = column_base(c.kind, c.index, a, scaled);
            if (!col) continue;
            // Sums that only feed a product stay lazy (limbs added, no
            // carries, no reduction): x + gamma < 2r with limbs below 2^30,
            // x + gamma + beta sigma < 3.1r and x + gamma + delta < 4r with
            // limbs below 2^31, each times a loose factor (< 2r) under the
            // product's r 2^261 limit, so the products stay loose.
            const S x = fr29_lazy_add(s_load(col + idx, scaled), gamma);
            const S sigma = s_load(a.perm_cosets + (size_t)(base + j) * len + idx, cs);
            left = s_mul(left, fr29_lazy_add(x, s_mul(beta, sigma)));
            right = s_mul(right, fr29_lazy_add(x, delta));
            if (base + j + 1 < a.perm_cols) delta = s_mul(delta, dconst);   // (the last column's is unused)
        }
        v = s_fold_diff(v, y, left, right, l_active);
    }
    return v;
}

// Lookup terms (numerator_lookup_kernel, bls12_381_g1_msm.cu:758-815;
// evaluation.rs:3405-3466).
template <bool kLds>
__device__ __forceinline__ S lookup_terms(S v, size_t idx, const NumArgs &a, const Slots<kLds> &sl)
{
    const size_t len = a.len;
    const size_t r_next = rotated(idx, 1, a.rot_scale, len);
    const size_t r_prev = rotated(idx, -1, a.rot_scale, len);
    const bool ks = a.scaled & kScLookup, ls = a.scaled & kScLagrange;
    const S one = s29_one();
    const S y = s_const(a.y), beta = s_const(a.beta), gamma = s_const(a.gamma);
    const S l0 = s_load(a.l0 + idx, ls);
    const S l_last = s_load(a.l_last + idx, ls);
    const S l_active = s_load(a.l_active + idx, ls);
    for (uint32_t n = 0; n < a.lookup_count; ++n) {
        const S table_value = eval_graph(a.graphs[1 + n], idx, a, sl);
        const Fr *z = a.lookup_products + (size_t)n * len;
        const Fr *ai = a.lookup_inputs + (size_t)n * len;
        const Fr *st = a.lookup_tables + (size_t)n * len;
        const S zi = s_load(z + idx, ks);
        const S a_i = s_load(ai + idx, ks);
        const S s_i = s_load(st + idx, ks);
        const S a_minus_s = s_sub(a_i, s_i);
        v = s_fold_diff(v, y, one, zi, l0);
        v = s_fold_diff(v, y, s_mul(zi, zi), zi, l_last);
        const S lhs = s_mul(s_mul(s_load(z + r_next, ks), s_add(a_i, beta)), s_add(s_i, gamma));
        v = s_fold_diff(v, y, lhs, s_mul(zi, table_value), l_active);
        v = s_fold(v, y, a_minus_s, l0);
        v = s_fold(v, y, s_mul(a_minus_s, s_sub(a_i, s_load(ai + r_prev, ks))), l_active);
    }
    return v;
}

template <bool kLds>
__device__ __forceinline__ Slots<kLds> slot_file(const NumArgs &a)
{
    if constexpr (kLds)
        return Slots<true>{threadIdx.x};
    else
        return Slots<false>{a.slots_global, a.slot_stride, blockIdx.x * (size_t)blockDim.x + threadIdx.x};
}

// kExtras = false: no lookup and no trash argument (a smaller kernel).
template <bool kLds, bool kExtras>
__global__ __launch_bounds__(kBlock, 4) void k_numerator(NumArgs a)
{
    const Slots<kLds> sl = slot_file<kLds>(a);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < a.len; idx += stride) {
        S v = a.run_custom ? eval_graph(a.graphs[0], idx, a, sl) : s_load(a.initial + idx, a.scaled & kScInitial);
        if (a.perm_sets != 0 && a.chunk_len != 0) v = permutation_terms(v, idx, a);
        if (kExtras) {
            if (a.lookup_count != 0) v = lookup_terms(v, idx, a, sl);
            const bool ts = a.scaled & kScTrash, fs = a.scaled & kScFixed;
            const S y = s_const(a.y), one = s29_one();
            for (uint32_t t = 0; t < a.trash_count; ++t) {
                const uint32_t sel = a.trash_sel[t];
                if (sel >= a.fixed_cols) continue;
                const S compressed = eval_graph(a.graphs[1 + a.lookup_count + t], idx, a, sl);
                const S q = s_load(a.fixed + (size_t)sel * a.len + idx, fs);
                const S trash = s_load(a.trash_cosets + (size_t)t * a.len + idx, ts);
                v = s_add(s_mul(v, y), s_sub(compressed, s_mul(s_sub(one, q), trash)));
            }
        }
        if (a.out_scaled) {
            fr29_reduce(v);
            store_fr(a.out + idx, fr29_to32(v));
        } else {
            store_fr(a.out + idx, s29_to_usual(v));
        }
    }
}

// dst[i] = src[i] * factor for the first n0 entries, src[i] after them (all
// scaled): the omega table with delta_start folded into its first level
__global__ __launch_bounds__(kBlock) void k_scale_prefix(Fr *dst, const Fr *src, size_t n0, size_t n, Fr factor)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    S v = fr29_from32(load_fr(src + i));
    if (i < n0) {
        v = s_mul(v, s_const(factor));
        fr29_reduce(v);
    }
    store_fr(dst + i, fr29_to32(v));
}

// v[i] (usual form) -> v[i] 2^5 mod r: the scaled form, stored as 8 x 32 bits
__global__ __launch_bounds__(kBlock) void k_to_scaled(Fr *v, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) store_fr(v + i, fr29_to32(s29_from_usual(load_fr(v + i))));
}

// out[g * len + idx] = graph g at row idx (eval_graph_batch, bls12_381_g1_msm.cu:5434-5515)
template <bool kLds>
__global__ __launch_bounds__(kBlock) void k_eval_graphs(NumArgs a)
{
    const Slots<kLds> sl = slot_file<kLds>(a);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < a.len; idx += stride)
        for (uint32_t gi = 0; gi < a.graph_count; ++gi)
            store_fr(a.out + (size_t)gi * a.len + idx, s29_to_usual(eval_graph(a.graphs[gi], idx, a, sl)));
}

// ---------------------------------------------------------------- host

// A graph copied to the device: one allocation holding constants,
// rotations, calculations and Horner parts.
struct UploadedGraph {
    DevBuf<uint8_t> blob;
    DevGraph g{};
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// a constant in the scaled form (v 2^261 mod r: the usual representation
// times 32), as the kernels take their challenges and graph constants
Fr scaled_fr(const HostFr &h)
{
    const HostFr x = h * HostFr::from_u64(32);
    Fr r;
    std::memcpy(r.v, x.l, 32);
    return r;
}
Fr scaled_fr(const RawFr &v) { return scaled_fr(to_hfr(v)); }

// omega^(j 2^(11 l)) tables (rowpow.hpp layout) in the scaled form
// (level 0 times `first` when given: delta_start folded in)
void build_scaled_omega_table(DevBuf<Fr> &tab, uint32_t &levels, const HostFr &omega, size_t len, hipStream_t st,
                              const HostFr *first = nullptr)
{
    uint32_t bits = 0;
    while (bits < 64 && (1ull << bits) < len) ++bits;
    levels = bits <= (uint32_t)kOmegaBits ? 1 : (bits + kOmegaBits - 1) / kOmegaBits;
    std::vector<HostFr> h((size_t)levels * kOmegaTab);
    const HostFr k32 = HostFr::from_u64(32);
    HostFr step = omega;
    for (uint32_t l = 0; l < levels; ++l) {
        HostFr acc = HostFr::one();
        const HostFr f = (l == 0 && first) ? k32 * *first : k32;
        for (uint32_t j = 0; j < kOmegaTab; ++j) {
            h[(size_t)l * kOmegaTab + j] = acc * f;
            acc = acc * step;
        }
        step = acc;   // omega^(2048 * 2^(11 l))
    }
    tab.alloc(h.size());
    MH_HIP(hipMemcpyAsync(tab.get(), h.data(), h.size() * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipStreamSynchronize(st));
}

void check_graph(const RawGraph &raw)
{
    MH_CHECK(raw.num_intermediates <= kMaxSlots, (int)hipErrorInvalidValue,
             "graph needs more than 256 intermediate slots");
    MH_CHECK((raw.constants_len == 0 || raw.constants) && (raw.rotations_len == 0 || raw.rotations) &&
                 (raw.calculations_len == 0 || raw.calculations) &&
                 (raw.horner_parts_len == 0 || raw.horner_parts),
             (int)hipErrorInvalidValue, "graph array is null");
    MH_CHECK(raw.calculations_len < (1ull << 32) && raw.constants_len < (1ull << 32) &&
                 raw.rotations_len < (1ull << 32) && raw.horner_parts_len < (1ull << 32),
             (int)hipErrorInvalidValue, "graph is too large");
    for (size_t c = 0; c < raw.calculations_len; ++c)
        MH_CHECK(raw.calculations[c].target < raw.num_intermediates, (int)hipErrorInvalidValue,
                 "graph calculation target is outside its intermediate slots");
}

void upload_graph(UploadedGraph &u, const RawGraph &raw, hipStream_t st)
{
    check_graph(raw);
    const size_t oc = 0;
    const size_t orot = oc + align256(raw.constants_len * sizeof(RawFr));
    const size_t ocalc = orot + align256(raw.rotations_len * sizeof(int32_t));
    const size_t opart = ocalc + align256(raw.calculations_len * sizeof(RawCalculation));
    const size_t total = opart + align256(raw.horner_parts_len * sizeof(RawValueSource));
    std::vector<uint8_t> host(total ? total : 1, 0);
    for (size_t i = 0; i < raw.constants_len; ++i) {   // constants in the scaled form
        const Fr c = scaled_fr(raw.constants[i]);
        std::memcpy(host.data() + oc + i * sizeof(RawFr), c.v, sizeof(RawFr));
    }
    if (raw.rotations_len) std::memcpy(host.data() + orot, raw.rotations, raw.rotations_len * sizeof(int32_t));
    if (raw.calculations_len)
        std::memcpy(host.data() + ocalc, raw.calculations, raw.calculations_len * sizeof(RawCalculation));
    if (raw.horner_parts_len)
        std::memcpy(host.data() + opart, raw.horner_parts, raw.horner_parts_len * sizeof(RawValueSource));
    u.blob.alloc(host.size());
    MH_HIP(hipMemcpyAsync(u.blob.get(), host.data(), host.size(), hipMemcpyHostToDevice, st));
    MH_HIP(hipStreamSynchronize(st));
    uint8_t *b = u.blob.get();
    u.g.constants = reinterpret_cast<const Fr *>(b + oc);
    u.g.rotations = reinterpret_cast<const int32_t *>(b + orot);
    u.g.calcs = reinterpret_cast<const RawCalculation *>(b + ocalc);
    u.g.parts = reinterpret_cast<const RawValueSource *>(b + opart);
    u.g.constants_len = (uint32_t)raw.constants_len;
    u.g.rotations_len = (uint32_t)raw.rotations_len;
    u.g.calcs_len = (uint32_t)raw.calculations_len;
    u.g.parts_len = (uint32_t)raw.horner_parts_len;
    u.g.num_slots = raw.num_intermediates;
}

// Device array of DevGraph descriptors, in evaluation order.
struct GraphSet {
    std::vector<std::unique_ptr<UploadedGraph>> graphs;
    DevBuf<DevGraph> table;
    uint32_t max_slots = 0;

    void add(const RawGraph &raw, hipStream_t st)
    {
        auto u = std::make_unique<UploadedGraph>();
        upload_graph(*u, raw, st);
        if (u->g.num_slots > max_slots) max_slots = u->g.num_slots;
        graphs.push_back(std::move(u));
    }
    void finish(hipStream_t st)
    {
        std::vector<DevGraph> h;
        for (auto &u : graphs) h.push_back(u->g);
        if (h.empty()) h.push_back(DevGraph{});
        table.alloc(h.size());
        MH_HIP(hipMemcpyAsync(table.get(), h.data(), h.size() * sizeof(DevGraph), hipMemcpyHostToDevice, st));
        MH_HIP(hipStreamSynchronize(st));
    }
};

template <class Kernel>
void launch_rows(NumArgs a, uint32_t slots, Kernel lds_kernel, Kernel global_kernel, hipStream_t st)
{
    if (a.len == 0) return;
    if (slots <= kLdsSlots) {
        const size_t blocks = ceil_div(a.len, kBlock);
        MH_CHECK(blocks <= 0x7fffffffu, (int)hipErrorInvalidValue, "too many rows");
        const size_t lds = (size_t)(slots ? slots : 1) * 36 * kBlock;
        hipLaunchKernelGGL(lds_kernel, dim3((unsigned)blocks), dim3(kBlock), lds, st, a);
    } else {
        const size_t blocks = std::min<size_t>(ceil_div(a.len, kBlock), kGlobalSlotBlocks);
        auto &sf = numerator_slot_file();
        a.slot_stride = blocks * kBlock;
        sf.buf.ensure((size_t)slots * 9 * a.slot_stride);
        a.slots_global = sf.buf.get();
        hipLaunchKernelGGL(global_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, st, a);
    }
    MH_HIP(hipGetLastError());
}

void run_numerator(const NumArgs &a, uint32_t slots, hipStream_t st)
{
    ProfileScope prof;
    prof.begin("numerator", st);
    if (a.lookup_count == 0 && a.trash_count == 0)
        launch_rows(a, slots, k_numerator<true, false>, k_numerator<false, false>, st);
    else
        launch_rows(a, slots, k_numerator<true, true>, k_numerator<false, true>, st);
    prof.end(st);
    if (profiling()) {
        MH_HIP(hipStreamSynchronize(st));
        prof.flush();
    }
}

// Host buffer -> device buffer (nullptr for empty ranges).
const Fr *up(DevBuf<Fr> &buf, const RawFr *src, size_t n, hipStream_t st)
{
    if (n == 0 || !src) return nullptr;
    buf.alloc(n);
    MH_HIP(hipMemcpyAsync(buf.get(), src, n * 32, hipMemcpyHostToDevice, st));
    return buf.get();
}

// A device buffer of usual-form values converted to the scaled form in place.
void to_scaled(Fr *v, size_t n, hipStream_t st)
{
    if (n == 0 || !v) return;
    hipLaunchKernelGGL(k_to_scaled, dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock), 0, st, v, n);
    MH_HIP(hipGetLastError());
}

// Host challenges -> device buffer in the scaled form.
const Fr *up_scaled(DevBuf<Fr> &buf, const RawFr *src, size_t n, hipStream_t st)
{
    if (n == 0 || !src) return nullptr;
    std::vector<Fr> h(n);
    for (size_t i = 0; i < n; ++i) h[i] = scaled_fr(src[i]);
    buf.alloc(n);
    MH_HIP(hipMemcpyAsync(buf.get(), h.data(), n * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipStreamSynchronize(st));   // h goes out of scope
    return buf.get();
}

template <class T>
const T *up_array(DevBuf<T> &buf, const T *src, size_t n, hipStream_t st)
{
    if (n == 0 || !src) return nullptr;
    buf.alloc(n);
    MH_HIP(hipMemcpyAsync(buf.get(), src, n * sizeof(T), hipMemcpyHostToDevice, st));
    return buf.get();
}

}  // namespace

// Global slot files are per lane: calls of one lane run in stream order.
LaneSlotFile &numerator_slot_file()
{
    static LaneSlotFile files[64][kLaneCount];
    return files[device_id()][current_lane()];
}

// Resident key data of one circuit (numerator_context_t,
// bls12_381_g1_msm.cu:1116-1160).
struct NumeratorContext {
    uint32_t lane = 0;
    size_t len = 0, input_len = 0, fixed_cols = 0, perm_cols = 0, perm_cosets_len = 0;
    size_t lookup_count = 0, trash_count = 0;
    int32_t rot_scale = 0, last_rotation = 0;
    uint32_t chunk_len = 0;
    HostFr omega{}, g_coset{}, delta_const{}, trace_omega{};
    // fixed, l0 / l_last / l_active and perm_cosets (extended cosets) are
    // held in the scaled form; fixed_lagrange and perm_lagrange (the trace
    // domain, read by the grand products and trace graphs) in the usual form
    DevBuf<Fr> fixed, fixed_lagrange, l0, l_last, l_active, perm_cosets, perm_lagrange, omega_tab;
    DevBuf<RawColumnRef> perm_columns;
    DevBuf<uint32_t> trash_sel;
    uint32_t omega_levels = 1;
    GraphSet graphs;                            // custom, lookups, trash
    GraphSet lookup_input_graphs, lookup_table_graphs;
};

}  // namespace mh

using namespace mh;

extern "C" RustError midnight_sppark_bls12_381_eval_numerator(
    RawFr *out, size_t len, const RawFr *initial_out, uint32_t run_custom, const RawFr *fixed,
    size_t fixed_cols, const RawFr *advice, size_t advice_cols, const RawFr *instance, size_t instance_cols,
    const RawFr *challenges, size_t challenge_len, RawFr beta, RawFr gamma, RawFr theta, RawFr trash_challenge,
    RawFr y, RawFr omega, RawFr delta_start, RawFr delta_const, int32_t rot_scale, int32_t last_rotation,
    uint32_t chunk_len, const RawFr *l0, const RawFr *l_last, const RawFr *l_active,
    const RawFr *permutation_products, size_t permutation_sets, const RawColumnRef *permutation_columns,
    size_t permutation_columns_len, const RawFr *permutation_pk_cosets, size_t permutation_pk_cosets_len,
    const RawFr *lookup_product_cosets, const RawFr *lookup_input_cosets, const RawFr *lookup_table_cosets,
    size_t lookup_count, const RawFr *trash_cosets, const uint32_t *trash_selector_fixed_indices,
    size_t trash_count, RawGraph custom_graph, const RawGraph *lookup_graphs, const RawGraph *trash_graphs)
{
    if (len == 0) return ok();
    if (!run_custom && initial_out == nullptr)
        return make_error((int)hipErrorInvalidValue, "initial_out is required when custom evaluation is skipped");
    if (!out || (fixed_cols && !fixed) || (advice_cols && !advice) || (instance_cols && !instance) ||
        (challenge_len && !challenges) ||
        ((permutation_sets || lookup_count) && (!l0 || !l_last || !l_active)) ||
        (permutation_sets && (!permutation_products || (permutation_columns_len && !permutation_columns) ||
                              (permutation_pk_cosets_len && !permutation_pk_cosets))) ||
        (lookup_count && (!lookup_graphs || !lookup_product_cosets || !lookup_input_cosets || !lookup_table_cosets)) ||
        (trash_count && (!trash_graphs || !trash_cosets || !trash_selector_fixed_indices)))
        return make_error((int)hipErrorInvalidValue, "invalid numerator arguments");
    if (permutation_sets && permutation_columns_len > permutation_pk_cosets_len)
        return make_error((int)hipErrorInvalidValue, "permutation columns and cosets do not match");
    if (fixed_cols > 0xffffffffu || advice_cols > 0xffffffffu || instance_cols > 0xffffffffu ||
        challenge_len > 0xffffffffu || permutation_sets > 0xffffffffu || permutation_columns_len > 0xffffffffu ||
        lookup_count > 0xffffffffu || trash_count > 0xffffffffu)
        return make_error((int)hipErrorInvalidValue, "numerator argument count too large");
    MH_TRY
    hipStream_t st = cur_stream();
    const size_t lag_len = (permutation_sets || lookup_count) ? len : 0;
    DevBuf<Fr> d_out(len), d_init, d_fixed, d_advice, d_instance, d_ch, d_l0, d_ll, d_la, d_pp, d_pc, d_lp, d_li,
        d_lt, d_tc, d_omega;
    DevBuf<RawColumnRef> d_cols;
    DevBuf<uint32_t> d_sel;
    NumArgs a{};
    a.out = d_out.get();
    a.len = len;
    a.run_custom = run_custom ? 1u : 0u;
    a.initial = run_custom ? nullptr : up(d_init, initial_out, len, st);
    a.fixed = up(d_fixed, fixed, fixed_cols * len, st);
    a.advice = up(d_advice, advice, advice_cols * len, st);
    a.instance = up(d_instance, instance, instance_cols * len, st);
    a.challenges = up_scaled(d_ch, challenges, challenge_len, st);
    a.fixed_cols = (uint32_t)fixed_cols;
    a.advice_cols = (uint32_t)advice_cols;
    a.instance_cols = (uint32_t)instance_cols;
    a.challenge_len = (uint32_t)challenge_len;
    a.beta = scaled_fr(beta);
    a.gamma = scaled_fr(gamma);
    a.theta = scaled_fr(theta);
    a.trash = scaled_fr(trash_challenge);
    a.y = scaled_fr(y);
    a.delta_start = scaled_fr(delta_start);
    a.delta_const = scaled_fr(delta_const);
    a.rot_scale = rot_scale;
    a.last_rotation = last_rotation;
    a.chunk_len = chunk_len;
    a.l0 = up(d_l0, l0, lag_len, st);
    a.l_last = up(d_ll, l_last, lag_len, st);
    a.l_active = up(d_la, l_active, lag_len, st);
    a.perm_sets = (uint32_t)permutation_sets;
    a.perm_cols = (uint32_t)permutation_columns_len;
    a.perm_products = up(d_pp, permutation_products, permutation_sets * len, st);
    a.perm_columns = up_array(d_cols, permutation_columns, permutation_columns_len, st);
    a.perm_cosets = up(d_pc, permutation_pk_cosets, permutation_pk_cosets_len * len, st);
    a.lookup_count = (uint32_t)lookup_count;
    a.lookup_products = up(d_lp, lookup_product_cosets, lookup_count * len, st);
    a.lookup_inputs = up(d_li, lookup_input_cosets, lookup_count * len, st);
    a.lookup_tables = up(d_lt, lookup_table_cosets, lookup_count * len, st);
    a.trash_count = (uint32_t)trash_count;
    a.trash_cosets = up(d_tc, trash_cosets, trash_count * len, st);
    a.trash_sel = up_array(d_sel, trash_selector_fixed_indices, trash_count, st);
    {
        const HostFr ds = to_hfr(delta_start);
        build_scaled_omega_table(d_omega, a.omega_levels, to_hfr(omega), len, st, &ds);
        a.omega_has_delta = 1;
    }
    a.omega_tab = d_omega.get();

    GraphSet gs;
    gs.add(run_custom ? custom_graph : RawGraph{}, st);
    for (size_t i = 0; i < lookup_count; ++i) gs.add(lookup_graphs[i], st);
    for (size_t i = 0; i < trash_count; ++i) gs.add(trash_graphs[i], st);
    gs.finish(st);
    a.graphs = gs.table.get();

    run_numerator(a, gs.max_slots, st);
    MH_HIP(hipMemcpyAsync(out, d_out.get(), len * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_eval_graph_batch(
    RawFr *out, size_t len, const RawFr *fixed, size_t fixed_cols, const RawFr *advice, size_t advice_cols,
    const RawFr *instance, size_t instance_cols, const RawFr *challenges, size_t challenge_len, RawFr beta,
    RawFr gamma, RawFr theta, RawFr trash_challenge, RawFr y, int32_t rot_scale, const RawGraph *graphs,
    size_t graph_count)
{
    if (len == 0 || graph_count == 0) return ok();
    if (!out || !graphs || (fixed_cols && !fixed) || (advice_cols && !advice) || (instance_cols && !instance) ||
        (challenge_len && !challenges))
        return make_error((int)hipErrorInvalidValue, "invalid graph batch arguments");
    if (fixed_cols > 0xffffffffu || advice_cols > 0xffffffffu || instance_cols > 0xffffffffu ||
        challenge_len > 0xffffffffu || graph_count > 0xffffffffu)
        return make_error((int)hipErrorInvalidValue, "graph batch argument count too large");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_out(len * graph_count), d_fixed, d_advice, d_instance, d_ch;
    NumArgs a{};
    a.out = d_out.get();
    a.len = len;
    a.fixed = up(d_fixed, fixed, fixed_cols * len, st);
    a.advice = up(d_advice, advice, advice_cols * len, st);
    a.instance = up(d_instance, instance, instance_cols * len, st);
    a.challenges = up_scaled(d_ch, challenges, challenge_len, st);
    a.fixed_cols = (uint32_t)fixed_cols;
    a.advice_cols = (uint32_t)advice_cols;
    a.instance_cols = (uint32_t)instance_cols;
    a.challenge_len = (uint32_t)challenge_len;
    a.beta = scaled_fr(beta);
    a.gamma = scaled_fr(gamma);
    a.theta = scaled_fr(theta);
    a.trash = scaled_fr(trash_challenge);
    a.y = scaled_fr(y);
    a.rot_scale = rot_scale;
    GraphSet gs;
    for (size_t i = 0; i < graph_count; ++i) gs.add(graphs[i], st);
    gs.finish(st);
    a.graphs = gs.table.get();
    a.graph_count = (uint32_t)graph_count;
    launch_rows(a, gs.max_slots, k_eval_graphs<true>, k_eval_graphs<false>, st);
    MH_HIP(hipMemcpyAsync(out, d_out.get(), len * graph_count * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_numerator_context_create(
    void **out, size_t len, size_t input_len, const RawFr *fixed, const RawFr *fixed_lagrange, size_t fixed_cols,
    const RawFr *l0, const RawFr *l_last, const RawFr *l_active, const RawColumnRef *permutation_columns,
    size_t permutation_columns_len, const RawFr *permutation_pk_cosets, const RawFr *permutation_pk_lagrange,
    size_t permutation_pk_cosets_len, const uint32_t *trash_selector_fixed_indices, size_t trash_count,
    RawGraph custom_graph, const RawGraph *lookup_graphs, const RawGraph *lookup_input_graphs,
    const RawGraph *lookup_table_graphs, size_t lookup_count, const RawGraph *trash_graphs, RawFr omega,
    RawFr trace_omega, RawFr g_coset, RawFr delta_const, int32_t rot_scale, int32_t last_rotation,
    uint32_t chunk_len)
{
    if (out == nullptr || len == 0 || input_len == 0 || l0 == nullptr || l_last == nullptr || l_active == nullptr)
        return make_error((int)hipErrorInvalidValue, "invalid numerator context arguments");
    if ((fixed_cols != 0 && (fixed == nullptr || fixed_lagrange == nullptr)) ||
        (permutation_columns_len != 0 && permutation_columns == nullptr) ||
        (permutation_pk_cosets_len != 0 && (permutation_pk_cosets == nullptr || permutation_pk_lagrange == nullptr)) ||
        (lookup_count != 0 &&
         (lookup_graphs == nullptr || lookup_input_graphs == nullptr || lookup_table_graphs == nullptr)) ||
        (trash_count != 0 && (trash_graphs == nullptr || trash_selector_fixed_indices == nullptr)))
        return make_error((int)hipErrorInvalidValue, "missing numerator context data");
    *out = nullptr;
    if (fixed_cols > 0xffffffffu || permutation_columns_len > 0xffffffffu || lookup_count > 0xffffffffu ||
        trash_count > 0xffffffffu || permutation_columns_len > permutation_pk_cosets_len)
        return make_error((int)hipErrorInvalidValue, "invalid numerator context arguments");
    MH_TRY
    hipStream_t st = cur_stream();
    auto c = std::make_unique<NumeratorContext>();
    c->lane = current_lane();
    c->len = len;
    c->input_len = input_len;
    c->fixed_cols = fixed_cols;
    c->perm_cols = permutation_columns_len;
    c->perm_cosets_len = permutation_pk_cosets_len;
    c->lookup_count = lookup_count;
    c->trash_count = trash_count;
    c->rot_scale = rot_scale;
    c->last_rotation = last_rotation;
    c->chunk_len = chunk_len;
    c->omega = to_hfr(omega);
    c->g_coset = to_hfr(g_coset);
    c->delta_const = to_hfr(delta_const);
    c->trace_omega = to_hfr(trace_omega);
    up(c->fixed, fixed, fixed_cols * len, st);
    up(c->fixed_lagrange, fixed_lagrange, fixed_cols * input_len, st);
    up(c->l0, l0, len, st);
    up(c->l_last, l_last, len, st);
    up(c->l_active, l_active, len, st);
    // the numerator reads the key's cosets in the scaled form (converted once here)
    to_scaled(c->fixed.get(), fixed_cols * len, st);
    to_scaled(c->l0.get(), len, st);
    to_scaled(c->l_last.get(), len, st);
    to_scaled(c->l_active.get(), len, st);
    up_array(c->perm_columns, permutation_columns, permutation_columns_len, st);
    up(c->perm_cosets, permutation_pk_cosets, permutation_pk_cosets_len * len, st);
    to_scaled(c->perm_cosets.get(), permutation_pk_cosets_len * len, st);
    up(c->perm_lagrange, permutation_pk_lagrange, permutation_pk_cosets_len * input_len, st);
    up_array(c->trash_sel, trash_selector_fixed_indices, trash_count, st);
    c->graphs.add(custom_graph, st);
    for (size_t i = 0; i < lookup_count; ++i) {
        c->graphs.add(lookup_graphs[i], st);
        c->lookup_input_graphs.add(lookup_input_graphs[i], st);
        c->lookup_table_graphs.add(lookup_table_graphs[i], st);
    }
    for (size_t i = 0; i < trash_count; ++i) c->graphs.add(trash_graphs[i], st);
    c->graphs.finish(st);
    c->lookup_input_graphs.finish(st);   // descriptor tables for trace_context_lookup_permute
    c->lookup_table_graphs.finish(st);
    build_scaled_omega_table(c->omega_tab, c->omega_levels, c->omega, len, st);
    MH_HIP(hipStreamSynchronize(st));
    *out = c.release();
    return ok();
    MH_CATCH
}

extern "C" void midnight_sppark_bls12_381_numerator_context_free(void *context)
{
    auto *c = reinterpret_cast<NumeratorContext *>(context);
    if (!c) return;
    try {
        (void)hipStreamSynchronize(lane_stream(c->lane));
    } catch (...) {
    }
    delete c;
}

namespace mh {

// numerator_context_eval; `columns_scaled` tells that the column vector holds
// its values in the scaled form (v 2^261 mod r) instead of the usual one (the
// proof pipeline's cosets, written so by the NTT that makes them).
RustError numerator_context_eval(void **out, const void *raw_context, const void *raw_columns, size_t total_columns,
                                 size_t advice_offset, size_t advice_cols, size_t instance_offset,
                                 size_t instance_cols, const RawFr *challenges, size_t challenge_len, RawFr beta,
                                 RawFr gamma, RawFr theta, RawFr trash_challenge, RawFr y, size_t permutation_offset,
                                 size_t permutation_sets, size_t lookup_product_offset, size_t lookup_input_offset,
                                 size_t lookup_table_offset, size_t lookup_count, size_t trash_offset,
                                 size_t trash_count, bool columns_scaled, bool out_scaled)
{
    if (out == nullptr || raw_context == nullptr || (challenge_len != 0 && challenges == nullptr))
        return make_error((int)hipErrorInvalidValue, "invalid numerator evaluation arguments");
    *out = nullptr;
    const auto *c = reinterpret_cast<const NumeratorContext *>(raw_context);
    const size_t len = c->len;
    const auto in_bounds = [total_columns](size_t offset, size_t count) {
        return offset <= total_columns && count <= total_columns - offset;
    };
    const auto *cols = reinterpret_cast<const DeviceVec *>(raw_columns);
    const bool vec_ok = total_columns == 0 ? (cols == nullptr || cols->len == 0)
                                           : (cols != nullptr && cols->len == total_columns * len);
    if (lookup_count != c->lookup_count || trash_count != c->trash_count || total_columns > SIZE_MAX / len ||
        !vec_ok || !in_bounds(advice_offset, advice_cols) || !in_bounds(instance_offset, instance_cols) ||
        !in_bounds(permutation_offset, permutation_sets) || !in_bounds(lookup_product_offset, lookup_count) ||
        !in_bounds(lookup_input_offset, lookup_count) || !in_bounds(lookup_table_offset, lookup_count) ||
        !in_bounds(trash_offset, trash_count))
        return make_error((int)hipErrorInvalidValue, "resident numerator vector length mismatch");
    if (advice_cols > 0xffffffffu || instance_cols > 0xffffffffu || challenge_len > 0xffffffffu ||
        permutation_sets > 0xffffffffu)
        return make_error((int)hipErrorInvalidValue, "numerator argument count too large");
    MH_TRY
    hipStream_t st = cur_stream();
    const Fr *base = cols ? cols->values.get() : nullptr;
    auto view = [base, len](size_t offset, size_t count) -> const Fr * {
        return count == 0 ? nullptr : base + offset * len;
    };
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = len;
    dv->values.alloc(len);
    DevBuf<Fr> d_ch;
    NumArgs a{};
    a.out = dv->values.get();
    a.len = len;
    a.run_custom = 1;
    a.fixed = c->fixed.get();
    a.fixed_cols = (uint32_t)c->fixed_cols;
    a.advice = view(advice_offset, advice_cols);
    a.advice_cols = (uint32_t)advice_cols;
    a.instance = view(instance_offset, instance_cols);
    a.instance_cols = (uint32_t)instance_cols;
    a.challenges = up_scaled(d_ch, challenges, challenge_len, st);
    a.challenge_len = (uint32_t)challenge_len;
    a.beta = scaled_fr(beta);
    a.gamma = scaled_fr(gamma);
    a.theta = scaled_fr(theta);
    a.trash = scaled_fr(trash_challenge);
    a.y = scaled_fr(y);
    const HostFr ds = to_hfr(beta) * c->g_coset;
    a.delta_start = scaled_fr(ds);
    a.delta_const = scaled_fr(c->delta_const);
    a.rot_scale = c->rot_scale;
    a.last_rotation = c->last_rotation;
    a.chunk_len = c->chunk_len;
    a.l0 = c->l0.get();
    a.l_last = c->l_last.get();
    a.l_active = c->l_active.get();
    a.perm_sets = (uint32_t)permutation_sets;
    a.perm_products = view(permutation_offset, permutation_sets);
    a.perm_columns = c->perm_columns.get();
    a.perm_cols = (uint32_t)c->perm_cols;
    a.perm_cosets = c->perm_cosets.get();
    // the context's omega table with this evaluation's delta_start folded
    // into its first level (one product per row fewer)
    DevBuf<Fr> d_omega(c->omega_levels * kOmegaTab);
    {
        const size_t n = (size_t)c->omega_levels * kOmegaTab;
        hipLaunchKernelGGL(k_scale_prefix, dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock), 0, st, d_omega.get(),
                           (const Fr *)c->omega_tab.get(), (size_t)kOmegaTab, n, a.delta_start);
        MH_HIP(hipGetLastError());
    }
    a.omega_tab = d_omega.get();
    a.omega_levels = c->omega_levels;
    a.omega_has_delta = 1;
    a.out_scaled = out_scaled ? 1u : 0u;
    a.lookup_count = (uint32_t)lookup_count;
    a.lookup_products = view(lookup_product_offset, lookup_count);
    a.lookup_inputs = view(lookup_input_offset, lookup_count);
    a.lookup_tables = view(lookup_table_offset, lookup_count);
    a.trash_count = (uint32_t)trash_count;
    a.trash_cosets = view(trash_offset, trash_count);
    a.trash_sel = c->trash_sel.get();
    a.graphs = c->graphs.table.get();
    a.scaled = kScFixed | kScCosets | kScLagrange |
               (columns_scaled ? (kScAdvice | kScInstance | kScProducts | kScLookup | kScTrash) : 0u);
    run_numerator(a, c->graphs.max_slots, st);
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}

}  // namespace mh

extern "C" RustError midnight_sppark_bls12_381_numerator_context_eval(
    void **out, const void *raw_context, const void *raw_columns, size_t total_columns, size_t advice_offset,
    size_t advice_cols, size_t instance_offset, size_t instance_cols, const RawFr *challenges, size_t challenge_len,
    RawFr beta, RawFr gamma, RawFr theta, RawFr trash_challenge, RawFr y, size_t permutation_offset,
    size_t permutation_sets, size_t lookup_product_offset, size_t lookup_input_offset, size_t lookup_table_offset,
    size_t lookup_count, size_t trash_offset, size_t trash_count)
{
    return numerator_context_eval(out, raw_context, raw_columns, total_columns, advice_offset, advice_cols,
                                  instance_offset, instance_cols, challenges, challenge_len, beta, gamma, theta,
                                  trash_challenge, y, permutation_offset, permutation_sets, lookup_product_offset,
                                  lookup_input_offset, lookup_table_offset, lookup_count, trash_offset, trash_count,
                                  false, false);
}

// bls12_381_g1_msm.cu:5667-5736: the context's trash graphs over its
// Lagrange domain (input_len rows, rotations unscaled), fixed columns in
// Lagrange form, no fold: out[t * input_len + row]
extern "C" RustError midnight_sppark_bls12_381_trace_context_trash(void **out, const void *raw_context,
                                                                   const void *raw_advice, size_t advice_cols,
                                                                   const void *raw_instance, size_t instance_cols,
                                                                   const RawFr *challenges, size_t challenge_len,
                                                                   RawFr trash_challenge)
{
    if (out == nullptr || raw_context == nullptr || (challenge_len != 0 && challenges == nullptr))
        return make_error((int)hipErrorInvalidValue, "invalid resident trash arguments");
    *out = nullptr;
    const auto *c = reinterpret_cast<const NumeratorContext *>(raw_context);
    const size_t len = c->input_len;
    const auto *advice = reinterpret_cast<const DeviceVec *>(raw_advice);
    const auto *instance = reinterpret_cast<const DeviceVec *>(raw_instance);
    auto has_len = [](const DeviceVec *v, size_t expected) {
        return expected == 0 ? (v == nullptr || v->len == 0) : (v != nullptr && v->len == expected);
    };
    if (c->trash_count == 0 || !has_len(advice, advice_cols * len) || !has_len(instance, instance_cols * len))
        return make_error((int)hipErrorInvalidValue, "resident trash vector length mismatch");
    if (advice_cols > 0xffffffffu || instance_cols > 0xffffffffu || challenge_len > 0xffffffffu)
        return make_error((int)hipErrorInvalidValue, "trash argument count too large");
    if (!c->graphs.table.get() || c->graphs.graphs.size() < 1 + c->lookup_count + c->trash_count)
        return make_error(-1, "numerator context holds no trash graphs");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_ch;
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = c->trash_count * len;
    dv->values.alloc(dv->len);
    NumArgs a{};
    a.out = dv->values.get();
    a.len = len;
    a.fixed = c->fixed_lagrange.get();
    a.advice = advice_cols ? advice->values.get() : nullptr;
    a.instance = instance_cols ? instance->values.get() : nullptr;
    a.challenges = up_scaled(d_ch, challenges, challenge_len, st);
    a.fixed_cols = (uint32_t)c->fixed_cols;
    a.advice_cols = (uint32_t)advice_cols;
    a.instance_cols = (uint32_t)instance_cols;
    a.challenge_len = (uint32_t)challenge_len;
    a.trash = scaled_fr(trash_challenge);
    a.rot_scale = 1;
    a.graphs = c->graphs.table.get() + 1 + c->lookup_count;   // custom, lookups, then trash
    a.graph_count = (uint32_t)c->trash_count;
    launch_rows(a, c->graphs.max_slots, k_eval_graphs<true>, k_eval_graphs<false>, st);
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:5849-5966: the lookups' compressed input and table
// columns from the context's lookup input / table graphs on its Lagrange
// domain (theta, rotations unscaled), then their permuted columns: a device
// vector [compressed inputs | compressed tables | permuted inputs | permuted
// tables], lookup-major blocks of input_len rows
extern "C" RustError midnight_sppark_bls12_381_trace_context_lookup_permute(
    void **out, const void *raw_context, const void *raw_advice, size_t advice_cols, const void *raw_instance,
    size_t instance_cols, const RawFr *challenges, size_t challenge_len, RawFr theta, size_t usable_rows,
    const RawFr *input_blinds, const RawFr *table_blinds, size_t blind_count)
{
    if (out == nullptr || raw_context == nullptr || (challenge_len != 0 && challenges == nullptr))
        return make_error((int)hipErrorInvalidValue, "invalid resident lookup permutation arguments");
    *out = nullptr;
    const auto *c = reinterpret_cast<const NumeratorContext *>(raw_context);
    const size_t len = c->input_len;
    const size_t lookup_count = c->lookup_count;
    const auto *advice = reinterpret_cast<const DeviceVec *>(raw_advice);
    const auto *instance = reinterpret_cast<const DeviceVec *>(raw_instance);
    auto has_len = [](const DeviceVec *v, size_t expected) {
        return expected == 0 ? (v == nullptr || v->len == 0) : (v != nullptr && v->len == expected);
    };
    if (lookup_count == 0 || usable_rows == 0 || usable_rows > len || blind_count != len - usable_rows ||
        c->lookup_input_graphs.graphs.size() != lookup_count || c->lookup_table_graphs.graphs.size() != lookup_count ||
        !has_len(advice, advice_cols * len) || !has_len(instance, instance_cols * len) ||
        (blind_count != 0 && (input_blinds == nullptr || table_blinds == nullptr)))
        return make_error((int)hipErrorInvalidValue, "resident lookup permutation shape mismatch");
    if (lookup_count > SIZE_MAX / len || lookup_count * len > SIZE_MAX / 4)
        return make_error((int)hipErrorInvalidValue, "resident lookup permutation size overflow");
    if (advice_cols > 0xffffffffu || instance_cols > 0xffffffffu || challenge_len > 0xffffffffu ||
        usable_rows > 0xffffffffu)
        return make_error((int)hipErrorInvalidValue, "lookup permutation argument count too large");
    if (!c->lookup_input_graphs.table.get() || !c->lookup_table_graphs.table.get() ||
        (!c->fixed_lagrange.get() && c->fixed_cols))
        return make_error(-1, "numerator context holds no lookup compression graphs");
    MH_TRY
    hipStream_t st = cur_stream();
    const size_t group = lookup_count * len;
    DevBuf<Fr> d_ch;
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = 4 * group;
    dv->values.alloc(dv->len);
    NumArgs a{};
    a.len = len;
    a.fixed = c->fixed_lagrange.get();
    a.advice = advice_cols ? advice->values.get() : nullptr;
    a.instance = instance_cols ? instance->values.get() : nullptr;
    a.challenges = up_scaled(d_ch, challenges, challenge_len, st);
    a.fixed_cols = (uint32_t)c->fixed_cols;
    a.advice_cols = (uint32_t)advice_cols;
    a.instance_cols = (uint32_t)instance_cols;
    a.challenge_len = (uint32_t)challenge_len;
    a.theta = scaled_fr(theta);
    a.rot_scale = 1;
    a.graph_count = (uint32_t)lookup_count;
    a.out = dv->values.get();
    a.graphs = c->lookup_input_graphs.table.get();
    launch_rows(a, c->lookup_input_graphs.max_slots, k_eval_graphs<true>, k_eval_graphs<false>, st);
    a.out = dv->values.get() + group;
    a.graphs = c->lookup_table_graphs.table.get();
    launch_rows(a, c->lookup_table_graphs.max_slots, k_eval_graphs<true>, k_eval_graphs<false>, st);
    DevBuf<Fr> blinds(std::max<size_t>(2 * lookup_count * blind_count, 1));
    if (blind_count) {
        MH_HIP(hipMemcpyAsync(blinds.get(), input_blinds, lookup_count * blind_count * 32, hipMemcpyHostToDevice, st));
        MH_HIP(hipMemcpyAsync(blinds.get() + lookup_count * blind_count, table_blinds, lookup_count * blind_count * 32,
                              hipMemcpyHostToDevice, st));
    }
    if (!lookup_permute_device(dv->values.get(), len, lookup_count, usable_rows, blinds.get(),
                               blinds.get() + lookup_count * blind_count, blind_count, st))
        return make_error(-1, "lookup input is not contained in the lookup table");
    *out = dv.release();
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_sppark_bls12_381_trace_context_permutation(
    void **out, const void *raw_context, const void *raw_advice, size_t advice_cols, const void *raw_instance,
    size_t instance_cols, RawFr beta, RawFr gamma, const RawFr *delta_powers, size_t delta_powers_len,
    const RawFr *blind_values, size_t blind_values_len, size_t blinding_factors)
{
    if (out == nullptr || raw_context == nullptr || delta_powers == nullptr)
        return make_error((int)hipErrorInvalidValue, "invalid resident permutation arguments");
    *out = nullptr;
    const auto *c = reinterpret_cast<const NumeratorContext *>(raw_context);
    const size_t len = c->input_len;
    const size_t column_count = c->perm_cols;
    const size_t chunk_len = c->chunk_len;
    const auto *advice = reinterpret_cast<const DeviceVec *>(raw_advice);
    const auto *instance = reinterpret_cast<const DeviceVec *>(raw_instance);
    auto has_len = [](const DeviceVec *v, size_t expected) {
        return expected == 0 ? (v == nullptr || v->len == 0) : (v != nullptr && v->len == expected);
    };
    if (column_count == 0 || chunk_len == 0 || delta_powers_len != column_count ||
        !has_len(advice, advice_cols * len) || !has_len(instance, instance_cols * len))
        return make_error((int)hipErrorInvalidValue, "resident permutation vector length mismatch");
    const size_t set_count = (column_count + chunk_len - 1) / chunk_len;
    if (blinding_factors >= len || blind_values_len != set_count * blinding_factors ||
        (blind_values_len != 0 && blind_values == nullptr))
        return make_error((int)hipErrorInvalidValue, "resident permutation blind mismatch");
    MH_TRY
    hipStream_t st = cur_stream();
    DevBuf<Fr> d_blinds;
    if (blind_values_len) {
        d_blinds.alloc(blind_values_len);
        MH_HIP(hipMemcpyAsync(d_blinds.get(), blind_values, blind_values_len * 32, hipMemcpyHostToDevice, st));
    }
    std::vector<HostFr> dp(column_count);
    for (size_t j = 0; j < column_count; ++j) dp[j] = to_hfr(delta_powers[j]);
    PermColumnsView v;
    v.fixed = c->fixed_lagrange.get();
    v.fixed_cols = (uint32_t)c->fixed_cols;
    v.advice = advice_cols ? advice->values.get() : nullptr;
    v.advice_cols = (uint32_t)advice_cols;
    v.instance = instance_cols ? instance->values.get() : nullptr;
    v.instance_cols = (uint32_t)instance_cols;
    v.refs = c->perm_columns.get();
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = set_count * len;
    dv->values.alloc(dv->len);
    permutation_products_device(dv->values.get(), len, column_count, chunk_len, v, c->perm_lagrange.get(),
                                dp.data(), to_hfr(beta), to_hfr(gamma), c->trace_omega, d_blinds.get(),
                                blinding_factors, st);
    *out = dv.release();
    return ok();
    MH_CATCH
}
