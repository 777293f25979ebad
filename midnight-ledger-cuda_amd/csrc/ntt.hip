// Fr NTT for gfx950: natural order in, natural order out.
//
// A transform of n = 2^log_n values is done as one or two LDS passes.
//   log_n <= 11: one pass; a workgroup holds the whole polynomial in LDS.
//   log_n <= 22: the four-step split n = n1 * n2 (n1 = 2^ceil(log_n/2)):
//     pass A, per column j2 (stride-n2 elements): an n1-point NTT, then the
//             element at row k1 is multiplied by omega_n^(j2*k1);
//     pass B, per row k1 (contiguous elements): an n2-point NTT whose output
//             k2 lands at index k1 + n1*k2.
//   X[k1 + n1 k2] = sum_j2 w_n2^(j2 k2) w_n^(j2 k1) sum_j1 x[j1 n2 + j2] w_n1^(j1 k1)
//   is the DFT, so the result equals the reference's bit_rev + CT NTT
//   (vendor/sppark-0.1.15/sppark/ntt/ntt.cuh:217-243) and best_fft
//   (vendor/midnight-curves-0.3.1/src/fft.rs:30-112) limb for limb.
// Inside a pass, a workgroup loads C columns of m = 2^t elements into LDS in
// bit-reversed order, runs t radix-2 DIT stages (in pairs, one barrier per
// pair; the first pair during the load, an odd last stage during the
// store), and writes the C columns back; C is chosen so C*m <= 2048.  Loads
// and stores walk the unit-stride dimension fastest, so every wave touches
// whole runs of C (pass A) or m (pass B) consecutive elements.
//
// Pre- and post-operations of the reference's fused entry points are folded
// into the first load and the last store of each transform:
//   pre  zeta: a_i *= {1, first, second}[i % 3], zero beyond input_len
//        (fr_coeff_to_extended_prepare_kernel, bls12_381_g1_msm.cu:1648-1678)
//   pre  repeat: a_i *= factors[i % factor_len]
//        (fr_mul_assign_repeating_kernel, :1213-1223)
//   post scale(+zeta): a_i *= scale * {1, first, second}[i % 3]
//        (fr_ntt_zeta_postprocess_kernel, :1680-1700)
//
// Arithmetic inside a pass is radix 2^29 (fr29.hpp): an element is repacked
// into 9 x 29-bit limbs at the load and stays so in LDS (36 bytes), still
// meaning a 2^256 mod r as in the usual form.  Twiddles and constants are
// stored as c 2^261 mod r, so the radix-2^29 Montgomery product a c 2^256
// 2^261 / 2^261 keeps the usual form.  Butterflies are lazy (Harvey): the
// twiddled operand comes out of the product below 2r, the sum and the
// difference (plus 4r) are formed limb by limb with no carries and no
// conditional subtraction, and the carries are propagated once per stage
// pair before the LDS store.  The first stage of a pass has only unit
// twiddles and skips its products, and so does the second stage's w_4^0
// half (its operand, below 4r, enters unmultiplied and is subtracted with
// 8r added): elements are below 2r at the load and below 12r after the
// first pair.  Every later
// stage multiplies, also by w^0 (in a stage's first butterfly), which keeps
// the bound growing by at most 4r per stage: below 48r after 11 stages,
// inside the 2^261 (70.6 r) the limbs hold and below what keeps every
// product's accumulators under 2^64.  The store reduces once
// to the canonical value and repacks to 8 x 32 bits, so results are limb
// for limb those of the reference.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "fields.hpp"
#include "fr29.hpp"
#include "host_fr.hpp"
#include "ntt.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

constexpr int kMaxLogLds = 11;     // largest single-pass sub-transform
constexpr int kTwLog = 22;         // twiddle tables are built for 2^22
constexpr int kThreads = 512;   // 8 waves: two 72 KiB workgroups per CU give 4 waves per SIMD
constexpr uint32_t kLdsElems = 2048;               // elements of a pass held in LDS
constexpr uint32_t kLdsElemBytes = 36;             // 9 x 29-bit limbs
constexpr int kLdsBytesMax = kLdsElems * kLdsElemBytes;

// A table of n entries held as three planes: limbs 0-3 and 4-7 as 16-byte
// words, limb 8 as a word, so an entry is two aligned 16-byte loads and one
// 4-byte load.
struct Tab29 {
    const uint4 *p0, *p1;
    const uint32_t *p2;
    __device__ __forceinline__ Fr29 get(uint32_t i) const
    {
        const uint4 a = p0[i], b = p1[i];
        Fr29 v;
        v.v[0] = a.x; v.v[1] = a.y; v.v[2] = a.z; v.v[3] = a.w;
        v.v[4] = b.x; v.v[5] = b.y; v.v[6] = b.z; v.v[7] = b.w;
        v.v[8] = p2[i];
        return v;
    }
};

struct Tables {
    // each entry c stored as c * 2^261 mod r in 29-bit limbs
    Tab29 roots[2];   // omega_{2^11}^j, j < 1024
    Tab29 lo[2];      // omega_{2^22}^i, i < 2048
    Tab29 hi[2];      // omega_{2^22}^(2048 i), i < 2048
};

// The tables live in static device memory ([direction][table] blocks of
// three planes, 32-bit words): filled once per device by a copy, so no
// transform ever allocates.
constexpr size_t kTwEntries = 1024 + 2048 + 2048;
__device__ __align__(256) uint32_t g_twiddles[2 * kTwEntries * 9];

Tables g_tables[64];
std::once_flag g_tables_once[64];

// c -> c * 2^261 mod r, which is the usual (2^256) form of 32 c
HostFr scaled29(const HostFr &c) { return c * HostFr::from_u64(32); }

// the 9 x 29-bit limbs of a host value's stored integer
void limbs29(const HostFr &h, uint32_t out[9])
{
    for (int j = 0; j < 9; ++j) {
        const int bit = 29 * j, q = bit >> 6, sh = bit & 63;
        uint64_t v = h.l[q] >> sh;
        if (sh > 35 && q + 1 < 4) v |= h.l[q + 1] << (64 - sh);
        out[j] = (uint32_t)v & 0x1fffffffu;
    }
}

Fr29 to_dev29(const HostFr &h)
{
    Fr29 r;
    limbs29(h, r.v);
    return r;
}

const Tables &tables()
{
    int dev = device_id();
    std::call_once(g_tables_once[dev], [dev]() {
        Tables &t = g_tables[dev];
        uint32_t *base = nullptr;
        MH_HIP(hipGetSymbolAddress(reinterpret_cast<void **>(&base), HIP_SYMBOL(g_twiddles)));
        std::vector<uint32_t> host(2 * kTwEntries * 9);
        // one table of n entries at word offset `off`: planes [n uint4][n uint4][n u32]
        auto put = [&](size_t off, size_t n, const std::vector<HostFr> &vals, Tab29 &tab) {
            uint32_t *w = host.data() + off;
            for (size_t i = 0; i < n; ++i) {
                uint32_t l[9];
                limbs29(scaled29(vals[i]), l);
                for (int j = 0; j < 4; ++j) {
                    w[4 * i + j] = l[j];
                    w[4 * n + 4 * i + j] = l[4 + j];
                }
                w[8 * n + i] = l[8];
            }
            tab.p0 = reinterpret_cast<const uint4 *>(base + off);
            tab.p1 = reinterpret_cast<const uint4 *>(base + off + 4 * n);
            tab.p2 = base + off + 8 * n;
        };
        for (int dir = 0; dir < 2; ++dir) {
            HostFr w22 = HostFr::omega(kTwLog);
            if (dir) w22 = w22.inverse();
            HostFr w11 = w22;
            for (int i = kMaxLogLds; i < kTwLog; ++i) w11 = w11.sqr();
            std::vector<HostFr> roots(1024), lo(2048), hi(2048);
            HostFr acc = HostFr::one();
            for (int i = 0; i < 1024; ++i) { roots[i] = acc; acc = acc * w11; }
            acc = HostFr::one();
            for (int i = 0; i < 2048; ++i) { lo[i] = acc; acc = acc * w22; }
            const HostFr step = acc;   // w22^2048
            acc = HostFr::one();
            for (int i = 0; i < 2048; ++i) { hi[i] = acc; acc = acc * step; }
            const size_t off = (size_t)dir * kTwEntries * 9;
            put(off, 1024, roots, t.roots[dir]);
            put(off + 1024 * 9, 2048, lo, t.lo[dir]);
            put(off + 3072 * 9, 2048, hi, t.hi[dir]);
        }
        MH_HIP(hipMemcpy(base, host.data(), host.size() * 4, hipMemcpyHostToDevice));
    });
    return g_tables[dev];
}

struct PassArgs {
    const Fr *in;
    Fr *out;
    uint64_t in_pstride, out_pstride;   // per polynomial (blockIdx.y)
    uint32_t in_es, in_cs;              // logical index = j*es + c*cs
    uint32_t out_es, out_cs;
    uint32_t log_m, log_C;              // column length and columns per workgroup (log2)
    uint32_t log_n;                     // whole transform (mid twiddle)
    uint32_t twiddle;                   // multiply (c, k) by w_n^(c*k)
    uint32_t pre;                       // 0 none, 1 zeta + zero pad, 2 repeating factors
    uint32_t post;                      // 0 none, 1 scale, 2 scale * zeta
    uint64_t in_len;                    // logical indices >= in_len read as zero
    const Fr *factors;
    uint64_t factor_len;
    Fr29 first, second, scale;          // stored as c * 2^261 mod r
    Tab29 roots, lo, hi;
    Tab29 mid;                          // w_n^(c k) at c * m + k (with 1/n for inverses), when has_mid
    uint32_t has_mid;
};

__device__ __forceinline__ uint32_t bitrev(uint32_t x, uint32_t bits)
{
    return __builtin_bitreverse32(x) >> (32 - bits);
}

// LDS holds a pass's C * m elements as three planes (limbs 0-3 and 4-7 as
// 16-byte words, limb 8 as a word), so lanes touching consecutive elements
// touch consecutive words of each plane: a wave's ds_read_b128 covers whole
// bank rows instead of colliding, as contiguous 36-byte elements would.
// Element e sits at slot e ^ (((e >> 3) ^ (e >> t)) & 7), t = log m - 3 (3
// for short columns, where the swizzle vanishes): the bit-reversed load
// (lanes 2^(log m - 3) elements apart) and the first stage pair (lanes four
// elements apart) then spread each group of eight lanes over the bank rows
// instead of stacking them on one.  The map permutes slots within aligned
// groups of eight, so it is a bijection on the pass's slots.
struct Lds29 {
    uint4 *p0, *p1;
    uint32_t *p2;
    uint32_t t;
    __device__ __forceinline__ Lds29(uint32_t *mem, uint32_t total, uint32_t log_m)
        : p0(reinterpret_cast<uint4 *>(mem)), p1(reinterpret_cast<uint4 *>(mem) + total),
          p2(mem + 8 * (size_t)total), t(log_m > 6 ? log_m - 3 : 3)
    {
    }
    __device__ __forceinline__ uint32_t slot(uint32_t e) const { return e ^ (((e >> 3) ^ (e >> t)) & 7u); }
    __device__ __forceinline__ Fr29 get(uint32_t e0) const
    {
        const uint32_t e = slot(e0);
        const uint4 a = p0[e], b = p1[e];
        Fr29 v;
        v.v[0] = a.x; v.v[1] = a.y; v.v[2] = a.z; v.v[3] = a.w;
        v.v[4] = b.x; v.v[5] = b.y; v.v[6] = b.z; v.v[7] = b.w;
        v.v[8] = p2[e];
        return v;
    }
    __device__ __forceinline__ void put(uint32_t e0, const Fr29 &v) const
    {
        const uint32_t e = slot(e0);
        p0[e] = make_uint4(v.v[0], v.v[1], v.v[2], v.v[3]);
        p1[e] = make_uint4(v.v[4], v.v[5], v.v[6], v.v[7]);
        p2[e] = v.v[8];
    }
};

__device__ __forceinline__ Fr29 zeta29(const Fr29 &v, uint64_t idx, const Fr29 &first, const Fr29 &second)
{
    const uint32_t c = (uint32_t)(idx % 3);
    if (c == 1) return fr29_mul_loose(v, first);
    if (c == 2) return fr29_mul_loose(v, second);
    return v;
}

// one workgroup of a pass: columns [bx * C, bx * C + C) of polynomial by
__device__ __forceinline__ void pass_body(const PassArgs &a, uint32_t bx, uint32_t by, uint32_t *lds_mem)
{
    const uint32_t log_m = a.log_m, m = 1u << log_m;
    const uint32_t log_C = a.log_C, C = 1u << log_C;
    const uint32_t total = C << log_m;
    const Lds29 lds(lds_mem, total, log_m);
    const uint32_t c0 = bx << log_C;
    const Fr *in = a.in + (uint64_t)by * a.in_pstride;
    Fr *out = a.out + (uint64_t)by * a.out_pstride;

    // an input element with the pre-operation: below 2r, limbs normalized
    // (zero at and past in_len)
    auto load_pre = [&](const Fr &x, uint64_t L) {
        Fr29 v;
        if (L < a.in_len) {
            Fr xv = x;
            if (a.pre == 2) xv = xv * a.factors[L % a.factor_len];
            v = fr29_from32(xv);
            if (a.pre == 1) v = zeta29(v, L, a.first, a.second);
        } else {
#pragma unroll
            for (int i = 0; i < 9; ++i) v.v[i] = 0;
        }
        return v;
    };

    // DIT stages, two at a time: a thread takes the four elements
    // e, e + h, e + 2h, e + 3h (h = 2^s, bits s and s+1 of e clear), does
    // stage s on (e, e+h) and (e+2h, e+3h) -- one twiddle w_2h^low for both
    // -- and stage s+1 on (e, e+2h) and (e+h, e+3h) -- twiddles w_4h^low and
    // w_4h^(low+h) -- with one LDS round trip and one barrier for the pair.
    // The first pair is fused with the load and an odd last stage (radix 2)
    // with the store.
    const bool load_cfast = (a.in_cs == 1) && C > 1;
    const uint32_t lq = log_m - 2;
    uint32_t s = 0;
    if (log_m >= 2) {
        // Load and first pair (h = 1, low = 0: only the product by w_4).
        // Group gg of a column holds LDS positions 4gg .. 4gg + 3, which the
        // bit-reversed order fills with the natural indices J, J + m/2,
        // J + m/4, J + 3m/4, J = bitrev(gg) over log m - 2 bits.  A thread
        // takes one J (lanes walk J, or the columns when those are
        // unit-stride), issues its four loads together, runs both stages in
        // registers and writes the group once.
        for (uint32_t f = threadIdx.x; f < (C << lq); f += blockDim.x) {
            uint32_t c, J;
            if (load_cfast) { c = f & (C - 1); J = f >> log_C; }
            else { J = f & ((1u << lq) - 1); c = f >> lq; }
            Fr x[4];
            uint64_t Ls[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t j = J + ((uint32_t)(((i & 1) << 1) | (i >> 1)) << lq);
                Ls[i] = (uint64_t)j * a.in_es + (uint64_t)(c0 + c) * a.in_cs;
                x[i] = Ls[i] < a.in_len ? load_fr(in + Ls[i]) : Fr{};
            }
            const Fr29 x0 = load_pre(x[0], Ls[0]), t1 = load_pre(x[1], Ls[1]);
            const Fr29 x2 = load_pre(x[2], Ls[2]), t3 = load_pre(x[3], Ls[3]);
            const Fr29 y0 = fr29_lazy_add(x0, t1), y1 = fr29_lazy_sub4r(x0, t1);
            const Fr29 y2 = fr29_lazy_add(x2, t3), y3 = fr29_lazy_sub4r(x2, t3);
            const Fr29 u3 = fr29_mul_loose(y3, a.roots.get(1u << (kMaxLogLds - 2)));
            // w_4^0 = 1: y2 (below 4r, limbs below 2^30) enters as it is
            Fr29 z0 = fr29_lazy_add(y0, y2), z2 = fr29_lazy_sub8r(y0, y2);
            Fr29 z1 = fr29_lazy_add(y1, u3), z3 = fr29_lazy_sub4r(y1, u3);
            fr29_norm(z0);
            fr29_norm(z1);
            fr29_norm(z2);
            fr29_norm(z3);
            const uint32_t e = (c << log_m) + ((lq ? bitrev(J, lq) : 0u) << 2);
            lds.put(e, z0);
            lds.put(e + 2, z2);
            lds.put(e + 1, z1);
            lds.put(e + 3, z3);
        }
        s = 2;
    } else {
        // one or two elements per column: load in bit-reversed order
        for (uint32_t f = threadIdx.x; f < total; f += blockDim.x) {
            uint32_t c, j;
            if (load_cfast) { c = f & (C - 1); j = f >> log_C; }
            else { j = f & (m - 1); c = f >> log_m; }
            const uint64_t L = (uint64_t)j * a.in_es + (uint64_t)(c0 + c) * a.in_cs;
            const Fr x = L < a.in_len ? load_fr(in + L) : Fr{};
            lds.put((c << log_m) + (log_m ? bitrev(j, log_m) : 0u), load_pre(x, L));
        }
    }
    __syncthreads();
    for (; s + 1 < log_m; s += 2) {
        const uint32_t h = 1u << s;
        for (uint32_t g = threadIdx.x; g < (C << lq); g += blockDim.x) {
            const uint32_t c = g >> lq, gg = g & ((1u << lq) - 1);
            const uint32_t low = gg & (h - 1);
            const uint32_t e = (c << log_m) + (((gg >> s) << (s + 2)) | low);
            const Fr29 x0 = lds.get(e), x2 = lds.get(e + 2 * h);
            Fr29 t1 = lds.get(e + h), t3 = lds.get(e + 3 * h);
            const Fr29 w = a.roots.get(low << (kMaxLogLds - 1 - s));
            t1 = fr29_mul_loose(t1, w);
            t3 = fr29_mul_loose(t3, w);
            const Fr29 y0 = fr29_lazy_add(x0, t1), y1 = fr29_lazy_sub4r(x0, t1);
            const Fr29 y2 = fr29_lazy_add(x2, t3), y3 = fr29_lazy_sub4r(x2, t3);
            const Fr29 u2 = fr29_mul_loose(y2, a.roots.get(low << (kMaxLogLds - 2 - s)));
            const Fr29 u3 = fr29_mul_loose(y3, a.roots.get((low + h) << (kMaxLogLds - 2 - s)));
            Fr29 z0 = fr29_lazy_add(y0, u2), z2 = fr29_lazy_sub4r(y0, u2);
            Fr29 z1 = fr29_lazy_add(y1, u3), z3 = fr29_lazy_sub4r(y1, u3);
            fr29_norm(z0);
            fr29_norm(z1);
            fr29_norm(z2);
            fr29_norm(z3);
            lds.put(e, z0);
            lds.put(e + 2 * h, z2);
            lds.put(e + h, z1);
            lds.put(e + 3 * h, z3);
        }
        __syncthreads();
    }

    // store, applying the mid twiddle or the post-operation, reducing to
    // the canonical value (elements are below 48r here; limbs below 2^31,
    // normalized unless they come from the fused last stage)
    const bool store_cfast = (a.out_cs == 1) && C > 1;
    const uint32_t tw_shift = kTwLog - a.log_n;
    auto store_one = [&](uint32_t c, uint32_t k, Fr29 v, bool lazy) {
        const uint64_t L = (uint64_t)k * a.out_es + (uint64_t)(c0 + c) * a.out_cs;
        if (a.twiddle) {
            // w_n^(c k) from the transform's table (below r), or formed from
            // the two small tables (w^0 = 1 included, below 1.06r); the
            // product is below 48r 1.06r / 2^261 + r < 1.8r
            Fr29 w;
            if (a.has_mid) {
                w = a.mid.get(((c0 + c) << log_m) + k);
            } else {
                const uint32_t e = ((c0 + c) * k) << tw_shift;
                w = a.lo.get(e & 2047);
                if (e >> 11) w = fr29_mul_loose(w, a.hi.get(e >> 11));
            }
            v = fr29_mul_loose(v, w);
            fr29_reduce(v);
        } else if (a.post == 3) {
            // the zeta pattern alone (below 1.1r after a factor, else < 48r)
            if (L % 3) {
                v = zeta29(v, L, a.first, a.second);
            } else {
                if (lazy) fr29_norm(v);
                fr29_reduce_top(v);
            }
            fr29_reduce(v);
        } else if (a.post) {
            // below 1.7r after the scale, below 1.1r after a zeta factor
            v = fr29_mul_loose(v, a.scale);
            if (a.post == 2) v = zeta29(v, L, a.first, a.second);
            fr29_reduce(v);
        } else {
            if (lazy) fr29_norm(v);
            fr29_reduce_top(v);
            fr29_reduce(v);
        }
        store_fr(out + L, fr29_to32(v));
    };
    if (s < log_m) {
        // last stage (radix 2, h = m/2) and store: butterfly k of a column
        // pairs k and k + m/2 (twiddle w_m^k, none when m = 2), and both
        // results go out from registers
        const uint32_t lh = log_m - 1, half = 1u << lh;
        for (uint32_t f = threadIdx.x; f < (C << lh); f += blockDim.x) {
            uint32_t c, k;
            if (store_cfast) { c = f & (C - 1); k = f >> log_C; }
            else { k = f & (half - 1); c = f >> lh; }
            const uint32_t e = (c << log_m) + k;
            const Fr29 u = lds.get(e);
            Fr29 v = lds.get(e + half);
            if (s) v = fr29_mul_loose(v, a.roots.get(k << (kMaxLogLds - 1 - s)));
            store_one(c, k, fr29_lazy_add(u, v), true);
            store_one(c, k + half, fr29_lazy_sub4r(u, v), true);
        }
    } else {
        for (uint32_t f = threadIdx.x; f < total; f += blockDim.x) {
            uint32_t c, k;
            if (store_cfast) { c = f & (C - 1); k = f >> log_C; }
            else { k = f & (m - 1); c = f >> log_m; }
            store_one(c, k, lds.get((c << log_m) + k), false);
        }
    }
}

// Workgroups are dealt to the 8 XCDs round robin by blockIdx.x; neighbouring
// column groups read the same 128-byte lines of a strided pass, so they are
// given to blocks on one XCD (sharing its L2): block b takes group
// (b % 8) * (gx / 8) + b / 8 when gx is a multiple of 8.
__global__ void __launch_bounds__(kThreads) ntt_pass(PassArgs a)
{
    extern __shared__ uint32_t lds[];
    const uint32_t b = blockIdx.x, gx = gridDim.x;
    const uint32_t bx = (gx & 7) ? b : (b & 7) * (gx >> 3) + (b >> 3);
    pass_body(a, bx, blockIdx.y, lds);
}

// A pass of every transform of a ragged set in one launch: descriptor d owns
// blocks [first_block, first_block + gx * npolys) of the grid.
struct RaggedDesc {
    PassArgs a;
    uint32_t gx, first_block;
};

__global__ void __launch_bounds__(kThreads) ntt_pass_ragged(const RaggedDesc *d, uint32_t nd)
{
    extern __shared__ uint32_t lds[];
    uint32_t lo = 0, hi = nd;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (d[mid].first_block <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const uint32_t local = blockIdx.x - d[lo].first_block;
    const uint32_t gx = d[lo].gx;
    const PassArgs a = d[lo].a;
    pass_body(a, local % gx, local / gx, lds);
}

// a 2^11-point column holds 72 KiB of LDS, above the default per-launch
// limit: raised once per device for both kernels
std::once_flag g_lds_once[64];
void lds_limit()
{
    std::call_once(g_lds_once[device_id()], []() {
        MH_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(ntt_pass),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytesMax));
        MH_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(ntt_pass_ragged),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytesMax));
    });
}

// log2 of the columns per workgroup: C * m <= 2048 elements (72 KiB)
uint32_t pick_log_cols(uint32_t log_m, uint32_t ncols)
{
    uint32_t lc = 0;
    while ((2u << lc) <= ncols && ((2u << lc) << log_m) <= kLdsElems) ++lc;
    return lc;
}

size_t lds_bytes(const PassArgs &a) { return ((size_t)1 << (a.log_C + a.log_m)) * kLdsElemBytes; }

void launch(const PassArgs &a, uint32_t ncols, uint32_t npolys, hipStream_t st)
{
    dim3 grid(ncols >> a.log_C, npolys);
    hipLaunchKernelGGL(ntt_pass, grid, dim3(kThreads), lds_bytes(a), st, a);
    MH_HIP(hipGetLastError());
}

// The four-step twiddles of a two-pass transform of 2^log_n: entry
// c * n1 + k = w_n^(c k) (c < n2 the column, k < n1 its row), times 1/n for
// the inverse direction, stored scaled in the three planes of Tab29.
__global__ void __launch_bounds__(256) k_mid_table(uint32_t log_n, uint32_t log_m, Tab29 lo, Tab29 hi, Fr29 factor,
                                                   uint32_t has_factor, uint4 *p0, uint4 *p1, uint32_t *p2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >> log_n) return;
    const uint32_t c = i >> log_m, k = i & ((1u << log_m) - 1u);
    const uint32_t e = (c * k) << (kTwLog - log_n);
    Fr29 w = lo.get(e & 2047);
    if (e >> 11) w = fr29_mul_loose(w, hi.get(e >> 11));
    if (has_factor) w = fr29_mul_loose(w, factor);
    fr29_reduce_2r(w);
    fr29_reduce(w);
    p0[i] = make_uint4(w.v[0], w.v[1], w.v[2], w.v[3]);
    p1[i] = make_uint4(w.v[4], w.v[5], w.v[6], w.v[7]);
    p2[i] = w.v[8];
}

// Tables per (device, log_n, direction), built on first use and kept until
// the process ends (2^22: 151 MB per direction).  A stream under graph
// capture gets none when the table is not built yet (the pass then forms
// its twiddles from the small tables): building allocates and waits.
struct MidEntry {
    bool built = false;
    Tab29 tab{};
};
std::mutex g_mid_mu;
MidEntry g_mid[64][2 * kMaxLogLds + 1][2];

const Tab29 *mid_table(uint32_t log_n, int dir, hipStream_t st)
{
    const int dev = device_id();
    std::lock_guard<std::mutex> g(g_mid_mu);
    MidEntry &m = g_mid[dev][log_n][dir];
    if (m.built) return &m.tab;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    MH_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return nullptr;
    const Tables &T = tables();
    const size_t n = size_t(1) << log_n;
    uint32_t *mem = nullptr;
    MH_HIP(hipMalloc(reinterpret_cast<void **>(&mem), n * kLdsElemBytes));
    uint4 *p0 = reinterpret_cast<uint4 *>(mem), *p1 = p0 + n;
    uint32_t *p2 = mem + 8 * n;
    const HostFr ninv = HostFr::from_u64(n).inverse();
    hipLaunchKernelGGL(k_mid_table, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, log_n,
                       (uint32_t)((log_n + 1) / 2), T.lo[dir], T.hi[dir], to_dev29(scaled29(ninv)), (uint32_t)dir,
                       p0, p1, p2);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(st));
    m.tab = Tab29{p0, p1, p2};
    m.built = true;
    return &m.tab;
}

}  // namespace

void ntt_warm_tables(uint32_t log_n, hipStream_t st)
{
    if (log_n <= kMaxLogLds || log_n > 2 * kMaxLogLds) return;
    for (int dir = 0; dir < 2; ++dir) (void)mid_table(log_n, dir, st);
}

namespace {

struct PassLaunch {
    PassArgs a;
    uint32_t ncols, npolys;
};

// the one (log n <= 11) or two passes of a job
int job_passes(const NttJob &job, PassLaunch out[2], hipStream_t st)
{
    const uint32_t log_n = job.log_n;
    MH_CHECK(log_n <= 2 * kMaxLogLds, hipErrorInvalidValue, "NTT size above 2^22 is not supported");
    const Tables &T = tables();
    lds_limit();
    const int dir = job.inverse ? 1 : 0;
    const uint64_t n = 1ull << log_n;
    // a two-pass transform reads its four-step twiddles (with 1/n for an
    // inverse) from its table when there is one
    const Tab29 *mid = (log_n > kMaxLogLds && job.mid_table) ? mid_table(log_n, dir, st) : nullptr;

    // post: inverse includes 1/n (unless the table holds it); ntt_zeta
    // multiplies by scale and the zeta pattern.  Factors equal to one are
    // not multiplied.
    const HostFr one = HostFr::one();
    HostFr post_scale = one;
    if (job.inverse && !mid) post_scale = HostFr::from_u64(n).inverse();
    if (job.post_zeta) post_scale = post_scale * job.scale;
    if (job.post_mul) post_scale = post_scale * job.mul;
    const bool post_pattern = job.post_zeta && !(job.first == one && job.second == one);
    const bool has_post = post_pattern || !(post_scale == one);
    const uint32_t post_mode = !has_post ? 0 : (post_scale == one ? 3 : (post_pattern ? 2 : 1));
    // the zeta pre-operation with unit factors is the zero padding alone
    const uint32_t pre = (job.pre == 1 && job.first == one && job.second == one) ? 0 : job.pre;

    PassArgs base{};
    base.roots = T.roots[dir];
    base.lo = T.lo[dir];
    base.hi = T.hi[dir];
    base.log_n = log_n;
    base.first = to_dev29(scaled29(job.first));
    base.second = to_dev29(scaled29(job.second));
    base.scale = to_dev29(scaled29(post_scale));
    base.factors = job.factors;
    base.factor_len = job.factor_len;

    if (log_n <= kMaxLogLds) {
        // (a length-1 transform is the identity; only pre/post factors apply)
        PassArgs a = base;
        a.in = job.in;
        a.out = job.out;
        a.in_pstride = job.in_pstride;
        a.out_pstride = job.out_pstride;
        a.in_es = 1; a.in_cs = 0;
        a.out_es = 1; a.out_cs = 0;
        a.log_m = log_n;
        a.log_C = 0;
        a.twiddle = 0;
        a.pre = pre;
        a.in_len = job.in_len;
        a.post = post_mode;
        out[0] = PassLaunch{a, 1, job.count};
        return 1;
    }

    const uint32_t t1 = (log_n + 1) / 2, t2 = log_n / 2;
    const uint32_t n1 = 1u << t1, n2 = 1u << t2;
    MH_CHECK(job.scratch != nullptr, hipErrorInvalidValue, "NTT scratch missing");

    PassArgs A = base;
    A.in = job.in;
    A.out = job.scratch;
    A.in_pstride = job.in_pstride;
    A.out_pstride = n;
    A.in_es = n2; A.in_cs = 1;
    A.out_es = n2; A.out_cs = 1;
    A.log_m = t1;
    A.log_C = pick_log_cols(t1, n2);
    A.twiddle = 1;
    A.pre = pre;
    A.in_len = job.in_len;
    A.post = 0;
    if (mid) {
        A.mid = *mid;
        A.has_mid = 1;
    }
    out[0] = PassLaunch{A, n2, job.count};

    PassArgs B = base;
    B.in = job.scratch;
    B.out = job.out;
    B.in_pstride = n;
    B.out_pstride = job.out_pstride;
    B.in_es = 1; B.in_cs = n2;
    B.out_es = n1; B.out_cs = 1;
    B.log_m = t2;
    B.log_C = pick_log_cols(t2, n1);
    B.twiddle = 0;
    B.pre = 0;
    B.in_len = ~0ull;
    B.post = post_mode;
    out[1] = PassLaunch{B, n1, job.count};
    return 2;
}

}  // namespace

void ntt_run(const NttJob &job, hipStream_t st)
{
    PassLaunch p[2];
    const int np = job_passes(job, p, st);
    for (int i = 0; i < np; ++i) launch(p[i].a, p[i].ncols, p[i].npolys, st);
}

// ---------------------------------------------------------------- ragged sets

struct RaggedNtt::Impl {
    DevBuf<RaggedDesc> descs[2];
    uint32_t nd[2] = {0, 0}, blocks[2] = {0, 0};
    size_t lds[2] = {0, 0};
    std::vector<RaggedDesc> host[2];
    DevBuf<Fr> scratch;
};

RaggedNtt::RaggedNtt() : impl_(new Impl) {}
RaggedNtt::~RaggedNtt() = default;

void RaggedNtt::prepare(Fr *base, const size_t *offsets, const uint32_t *log_ns, size_t count, bool inverse,
                        hipStream_t st)
{
    Impl &im = *impl_;
    // runs of consecutive, back-to-back transforms of one size share a
    // descriptor (one grid row per polynomial)
    struct Run {
        size_t first, count;
    };
    std::vector<Run> runs;
    size_t scratch = 0;
    for (size_t i = 0; i < count;) {
        size_t j = i + 1;
        const size_t n = size_t(1) << log_ns[i];
        while (j < count && log_ns[j] == log_ns[i] && offsets[j] == offsets[j - 1] + n) ++j;
        runs.push_back(Run{i, j - i});
        scratch += ntt_scratch_elems(log_ns[i], j - i);
        i = j;
    }
    im.scratch.ensure(std::max<size_t>(scratch, 1));
    for (int s = 0; s < 2; ++s) {
        im.host[s].clear();
        im.blocks[s] = 0;
        im.lds[s] = 0;
    }
    size_t soff = 0;
    for (const Run &r : runs) {
        const uint32_t log_n = log_ns[r.first];
        NttJob job;
        job.in = base + offsets[r.first];
        job.out = base + offsets[r.first];
        job.in_pstride = job.out_pstride = uint64_t(1) << log_n;
        job.log_n = log_n;
        job.count = (uint32_t)r.count;
        job.inverse = inverse;
        job.scratch = im.scratch.get() + soff;
        soff += ntt_scratch_elems(log_n, r.count);
        PassLaunch p[2];
        const int np = job_passes(job, p, st);
        for (int s = 0; s < np; ++s) {
            const uint32_t gx = p[s].ncols >> p[s].a.log_C;
            const uint64_t nb = (uint64_t)gx * p[s].npolys;
            MH_CHECK((uint64_t)im.blocks[s] + nb < (1ull << 31), hipErrorInvalidValue, "ragged NTT set too large");
            im.host[s].push_back(RaggedDesc{p[s].a, gx, im.blocks[s]});
            im.blocks[s] += (uint32_t)nb;
            im.lds[s] = std::max<size_t>(im.lds[s], lds_bytes(p[s].a));
        }
    }
    for (int s = 0; s < 2; ++s) {
        im.nd[s] = (uint32_t)im.host[s].size();
        if (!im.nd[s]) continue;
        im.descs[s].ensure(im.nd[s]);
        MH_HIP(hipMemcpyAsync(im.descs[s].get(), im.host[s].data(), im.nd[s] * sizeof(RaggedDesc),
                              hipMemcpyHostToDevice, st));
    }
}

void RaggedNtt::enqueue(hipStream_t st)
{
    const Impl &im = *impl_;
    for (int s = 0; s < 2; ++s) {
        if (!im.nd[s]) continue;
        hipLaunchKernelGGL(ntt_pass_ragged, dim3(im.blocks[s]), dim3(kThreads), im.lds[s], st,
                           (const RaggedDesc *)im.descs[s].get(), im.nd[s]);
        MH_HIP(hipGetLastError());
    }
}

size_t ntt_scratch_elems(uint32_t log_n, size_t count)
{
    return log_n <= kMaxLogLds ? 0 : (size_t(1) << log_n) * count;
}

}  // namespace mh
