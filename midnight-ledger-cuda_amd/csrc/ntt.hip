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
// bit-reversed order, runs t radix-2 DIT stages with one barrier per stage,
// and writes the C columns back; C is chosen so C*m*32 B <= 64 KiB.  Loads
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
// Products of a data value by a constant (twiddles, zeta, scale) use the
// radix-2^29 product (fr29.hpp, ~0.6x the instructions of the 32-bit one):
// the twiddle tables and the constants are stored as c 2^261 mod r, so the
// result comes out in the usual 2^256 form, limb-identical.  The one
// data-by-data product (pre-multiplication by caller factors) stays 32-bit.
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
constexpr int kThreads = 512;   // 8 waves: two 64 KiB workgroups per CU give 4 waves per SIMD

struct Tables {
    // each stored as value * 2^261 mod r (the radix-2^29 product's form)
    Fr *roots[2];   // omega_{2^11}^j, j < 1024
    Fr *lo[2];      // omega_{2^22}^i, i < 2048
    Fr *hi[2];      // omega_{2^22}^(2048 i), i < 2048
};

// The tables live in static device memory ([direction][table][entry], 32-bit
// words): filled once per device by a copy, so no transform ever allocates.
constexpr size_t kTwEntries = 1024 + 2048 + 2048;
__device__ __align__(256) uint32_t g_twiddles[2 * kTwEntries * 8];

Tables g_tables[64];
std::once_flag g_tables_once[64];

// c -> c * 2^261 mod r, which is the usual (2^256) form of 32 c
HostFr scaled29(const HostFr &c) { return c * HostFr::from_u64(32); }

const Tables &tables()
{
    int dev = device_id();
    std::call_once(g_tables_once[dev], [dev]() {
        Tables &t = g_tables[dev];
        Fr *base = nullptr;
        MH_HIP(hipGetSymbolAddress(reinterpret_cast<void **>(&base), HIP_SYMBOL(g_twiddles)));
        std::vector<HostFr> host(2 * kTwEntries);
        for (int dir = 0; dir < 2; ++dir) {
            HostFr w22 = HostFr::omega(kTwLog);
            if (dir) w22 = w22.inverse();
            HostFr w11 = w22;
            for (int i = kMaxLogLds; i < kTwLog; ++i) w11 = w11.sqr();
            HostFr *roots = host.data() + dir * kTwEntries, *lo = roots + 1024, *hi = lo + 2048;
            HostFr acc = HostFr::one();
            for (int i = 0; i < 1024; ++i) { roots[i] = scaled29(acc); acc = acc * w11; }
            acc = HostFr::one();
            for (int i = 0; i < 2048; ++i) { lo[i] = scaled29(acc); acc = acc * w22; }
            HostFr step = acc;   // w22^2048
            acc = HostFr::one();
            for (int i = 0; i < 2048; ++i) { hi[i] = scaled29(acc); acc = acc * step; }
            t.roots[dir] = base + dir * kTwEntries;
            t.lo[dir] = t.roots[dir] + 1024;
            t.hi[dir] = t.lo[dir] + 2048;
        }
        MH_HIP(hipMemcpy(base, host.data(), host.size() * 32, hipMemcpyHostToDevice));
    });
    return g_tables[dev];
}

struct PassArgs {
    const Fr *in;
    Fr *out;
    uint64_t in_pstride, out_pstride;   // per polynomial (blockIdx.y)
    uint32_t in_es, in_cs;              // logical index = j*es + c*cs
    uint32_t out_es, out_cs;
    uint32_t log_m, C;
    uint32_t log_n;                     // whole transform (mid twiddle)
    uint32_t twiddle;                   // multiply (c, k) by w_n^(c*k)
    uint32_t pre;                       // 0 none, 1 zeta + zero pad, 2 repeating factors
    uint32_t post;                      // 0 none, 1 scale, 2 scale * zeta
    uint64_t in_len;                    // logical indices >= in_len read as zero
    const Fr *factors;
    uint64_t factor_len;
    Fr first, second, scale;            // stored as c * 2^261 mod r
    const Fr *roots, *lo, *hi;
};

__device__ __forceinline__ uint32_t bitrev(uint32_t x, uint32_t bits)
{
    return __builtin_bitreverse32(x) >> (32 - bits);
}

__device__ __forceinline__ Fr zeta_factor(const Fr &v, uint64_t idx, const Fr &first, const Fr &second)
{
    uint32_t c = (uint32_t)(idx % 3);
    if (c == 1) return fr_mul_scaled(v, first);
    if (c == 2) return fr_mul_scaled(v, second);
    return v;
}

// LDS holds a pass's C * m elements as two planes of 16-byte halves
// (element e: low half at plane[e], high half at plane[P + e]), so lanes
// touching consecutive elements touch consecutive 16-byte words: a wave's
// ds_read_b128 then covers whole bank rows instead of colliding 8 ways, as
// contiguous 32-byte elements do.
struct LdsPlanes {
    uint4 *plane;
    uint32_t P;
    __device__ __forceinline__ Fr get(uint32_t e) const
    {
        const uint4 lo = plane[e], hi = plane[P + e];
        Fr v;
        v.v[0] = lo.x; v.v[1] = lo.y; v.v[2] = lo.z; v.v[3] = lo.w;
        v.v[4] = hi.x; v.v[5] = hi.y; v.v[6] = hi.z; v.v[7] = hi.w;
        return v;
    }
    __device__ __forceinline__ void put(uint32_t e, const Fr &v) const
    {
        plane[e] = make_uint4(v.v[0], v.v[1], v.v[2], v.v[3]);
        plane[P + e] = make_uint4(v.v[4], v.v[5], v.v[6], v.v[7]);
    }
};

// one workgroup of a pass: columns [bx * C, bx * C + C) of polynomial by
__device__ __forceinline__ void pass_body(const PassArgs &a, uint32_t bx, uint32_t by, Fr *lds_mem)
{
    const uint32_t m = 1u << a.log_m;
    const uint32_t C = a.C;
    const uint32_t total = C * m;
    const LdsPlanes lds{reinterpret_cast<uint4 *>(lds_mem), total};
    const uint32_t c0 = bx * C;
    const Fr *in = a.in + (uint64_t)by * a.in_pstride;
    Fr *out = a.out + (uint64_t)by * a.out_pstride;

    // load (bit-reversed within each column), applying the pre-operation
    const bool load_cfast = (a.in_cs == 1) && C > 1;
    for (uint32_t f = threadIdx.x; f < total; f += blockDim.x) {
        uint32_t c, j;
        if (load_cfast) { c = f % C; j = f / C; }
        else { j = f % m; c = f / m; }
        uint64_t L = (uint64_t)j * a.in_es + (uint64_t)(c0 + c) * a.in_cs;
        Fr v;
        if (L < a.in_len) {
            v = load_fr(in + L);
            if (a.pre == 1) v = zeta_factor(v, L, a.first, a.second);
            else if (a.pre == 2) v = v * a.factors[L % a.factor_len];
        } else {
            v = Fr::zero();
        }
        lds.put(c * m + bitrev(j, a.log_m), v);
    }
    __syncthreads();

    // DIT stages, two at a time: a thread takes the four elements
    // e, e + h, e + 2h, e + 3h (h = 2^s, bits s and s+1 of e clear), does
    // stage s on (e, e+h) and (e+2h, e+3h) -- one twiddle w_2h^low for both
    // -- and stage s+1 on (e, e+2h) and (e+h, e+3h) -- twiddles w_4h^low and
    // w_4h^(low+h) -- with one LDS round trip and one barrier for the pair.
    // An odd last stage is a radix-2 one.
    uint32_t s = 0;
    for (; s + 1 < a.log_m; s += 2) {
        const uint32_t h = 1u << s;
        const uint32_t quarter = m >> 2;
        for (uint32_t g = threadIdx.x; g < C * quarter; g += blockDim.x) {
            const uint32_t c = g / quarter, gg = g % quarter;
            const uint32_t low = gg & (h - 1);
            const uint32_t e = c * m + (((gg >> s) << (s + 2)) | low);
            Fr x0 = lds.get(e), x1 = lds.get(e + h), x2 = lds.get(e + 2 * h), x3 = lds.get(e + 3 * h);
            if (low) {
                const Fr29 w = fr29_from32(a.roots[low << (kMaxLogLds - 1 - s)]);
                x1 = fr_mul_scaled29(x1, w);
                x3 = fr_mul_scaled29(x3, w);
            }
            const Fr y0 = x0 + x1, y1 = x0 - x1;
            Fr y2 = x2 + x3, y3 = x2 - x3;
            if (low) y2 = fr_mul_scaled(y2, a.roots[low << (kMaxLogLds - 2 - s)]);
            y3 = fr_mul_scaled(y3, a.roots[(low + h) << (kMaxLogLds - 2 - s)]);
            lds.put(e, y0 + y2);
            lds.put(e + 2 * h, y0 - y2);
            lds.put(e + h, y1 + y3);
            lds.put(e + 3 * h, y1 - y3);
        }
        __syncthreads();
    }
    if (s < a.log_m) {
        const uint32_t h = 1u << s;
        const uint32_t half = m >> 1;
        for (uint32_t b = threadIdx.x; b < C * half; b += blockDim.x) {
            const uint32_t c = b / half, bb = b % half;
            const uint32_t low = bb & (h - 1);
            const uint32_t e = c * m + (((bb >> s) << (s + 1)) | low);
            Fr u = lds.get(e);
            Fr v = lds.get(e + h);
            if (low) v = fr_mul_scaled(v, a.roots[low << (kMaxLogLds - 1 - s)]);
            lds.put(e, u + v);
            lds.put(e + h, u - v);
        }
        __syncthreads();
    }

    // store, applying the mid twiddle and the post-operation
    const bool store_cfast = (a.out_cs == 1) && C > 1;
    const uint32_t tw_shift = kTwLog - a.log_n;
    for (uint32_t f = threadIdx.x; f < total; f += blockDim.x) {
        uint32_t c, k;
        if (store_cfast) { c = f % C; k = f / C; }
        else { k = f % m; c = f / m; }
        Fr v = lds.get(c * m + k);
        if (a.twiddle) {
            uint32_t e = ((c0 + c) * k) << tw_shift;
            if (e) {
                Fr w = a.lo[e & 2047];
                if (e >> 11) w = fr_mul_scaled(w, a.hi[e >> 11]);   // stays in the 2^261 form
                v = fr_mul_scaled(v, w);
            }
        }
        uint64_t L = (uint64_t)k * a.out_es + (uint64_t)(c0 + c) * a.out_cs;
        if (a.post) {
            v = fr_mul_scaled(v, a.scale);
            if (a.post == 2) v = zeta_factor(v, L, a.first, a.second);
        }
        store_fr(out + L, v);
    }
}

__global__ void __launch_bounds__(kThreads) ntt_pass(PassArgs a)
{
    extern __shared__ Fr lds[];
    pass_body(a, blockIdx.x, blockIdx.y, lds);
}

// A pass of every transform of a ragged set in one launch: descriptor d owns
// blocks [first_block, first_block + gx * npolys) of the grid.
struct RaggedDesc {
    PassArgs a;
    uint32_t gx, first_block;
};

__global__ void __launch_bounds__(kThreads) ntt_pass_ragged(const RaggedDesc *d, uint32_t nd)
{
    extern __shared__ Fr lds[];
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

uint32_t pick_cols(uint32_t log_m, uint32_t ncols)
{
    uint32_t C = 1;
    while (C * 2 <= ncols && ((C * 2) << log_m) <= 2048) C *= 2;
    return C;
}

void launch(const PassArgs &a, uint32_t ncols, uint32_t npolys, hipStream_t st)
{
    dim3 grid(ncols / a.C, npolys);
    size_t lds = (size_t)a.C * (1u << a.log_m) * sizeof(Fr);
    hipLaunchKernelGGL(ntt_pass, grid, dim3(kThreads), lds, st, a);
    MH_HIP(hipGetLastError());
}

Fr to_dev(const HostFr &h)
{
    Fr r;
    std::memcpy(r.v, h.l, 32);
    return r;
}

}  // namespace

namespace {

struct PassLaunch {
    PassArgs a;
    uint32_t ncols, npolys;
};

// the one (log n <= 11) or two passes of a job
int job_passes(const NttJob &job, PassLaunch out[2])
{
    const uint32_t log_n = job.log_n;
    MH_CHECK(log_n <= 2 * kMaxLogLds, hipErrorInvalidValue, "NTT size above 2^22 is not supported");
    const Tables &T = tables();
    const int dir = job.inverse ? 1 : 0;
    const uint64_t n = 1ull << log_n;

    // post: inverse includes 1/n; ntt_zeta multiplies by scale and the zeta pattern
    HostFr post_scale = HostFr::one();
    bool has_post = false;
    if (job.inverse) {
        post_scale = HostFr::from_u64(n).inverse();
        has_post = true;
    }
    if (job.post_zeta) {
        post_scale = post_scale * job.scale;
        has_post = true;
    }
    if (job.post_mul) {
        post_scale = post_scale * job.mul;
        has_post = true;
    }

    PassArgs base{};
    base.roots = T.roots[dir];
    base.lo = T.lo[dir];
    base.hi = T.hi[dir];
    base.log_n = log_n;
    base.first = to_dev(scaled29(job.first));
    base.second = to_dev(scaled29(job.second));
    base.scale = to_dev(scaled29(post_scale));
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
        a.C = 1;
        a.twiddle = 0;
        a.pre = job.pre;
        a.in_len = job.in_len;
        a.post = has_post ? (job.post_zeta ? 2 : 1) : 0;
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
    A.C = pick_cols(t1, n2);
    A.twiddle = 1;
    A.pre = job.pre;
    A.in_len = job.in_len;
    A.post = 0;
    out[0] = PassLaunch{A, n2, job.count};

    PassArgs B = base;
    B.in = job.scratch;
    B.out = job.out;
    B.in_pstride = n;
    B.out_pstride = job.out_pstride;
    B.in_es = 1; B.in_cs = n2;
    B.out_es = n1; B.out_cs = 1;
    B.log_m = t2;
    B.C = pick_cols(t2, n1);
    B.twiddle = 0;
    B.pre = 0;
    B.in_len = ~0ull;
    B.post = has_post ? (job.post_zeta ? 2 : 1) : 0;
    out[1] = PassLaunch{B, n1, job.count};
    return 2;
}

}  // namespace

void ntt_run(const NttJob &job, hipStream_t st)
{
    PassLaunch p[2];
    const int np = job_passes(job, p);
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
        const int np = job_passes(job, p);
        for (int s = 0; s < np; ++s) {
            const uint32_t gx = p[s].ncols / p[s].a.C;
            const uint64_t nb = (uint64_t)gx * p[s].npolys;
            MH_CHECK((uint64_t)im.blocks[s] + nb < (1ull << 31), hipErrorInvalidValue, "ragged NTT set too large");
            im.host[s].push_back(RaggedDesc{p[s].a, gx, im.blocks[s]});
            im.blocks[s] += (uint32_t)nb;
            im.lds[s] = std::max<size_t>(im.lds[s], (size_t)p[s].a.C * (1u << p[s].a.log_m) * sizeof(Fr));
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
