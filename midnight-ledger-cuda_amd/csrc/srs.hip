// KZG parameter files (ParamsKZG::write_custom / read_custom,
// vendor/midnight-proofs-0.8.1/src/poly/kzg/params.rs:170-262) parsed on the
// device into resident MSM contexts, so that the prover keeps the SRS in
// HBM without a host round trip of 2 x 2^k points:
//
//   u32 k (LE) || 2^k G1 g ([tau^i]G) || 2^k G1 g_lagrange || G2 || [tau]G2
//
// Formats (utils/helpers.rs:16-28 SerdeFormat, in its order):
//   0 Processed:          G1 compressed (48 B, ZCash flags), decompressed with
//                         a square root, subgroup-checked as from_bytes is;
//   1 RawBytes:           G1 uncompressed (96 B: x || y big-endian, ZCash
//                         flags), coordinates below p, on the curve and in the
//                         r-torsion subgroup (G1Affine::read_raw, g1.rs:469-481);
//   2 RawBytesUnchecked:  the same bytes, coordinates below p, no curve or
//                         subgroup check (read_raw_unchecked).
// The G2 elements are carried as the bytes read (the prover does not use
// them).  One lane per point: flags, big-endian bytes to 12 x 32-bit limbs,
// canonical -> Montgomery, the checks, and the RawG1Affine row (identity all
// zero) written straight into the context's point rows.
#include <cstring>
#include <fstream>
#include <memory>
#include <vector>

#include "curve.hpp"
#include "fields.hpp"
#include "msm_context.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

constexpr uint32_t kProcessed = 0, kRawBytes = 1, kRawBytesUnchecked = 2;

__device__ __forceinline__ Fp fp_const(const uint32_t (&c)[12])
{
    Fp r;
#pragma unroll
    for (int i = 0; i < 12; ++i) r.v[i] = c[i];
    return r;
}

// R^2 mod p (turns a canonical value into Montgomery form by one product)
__device__ __forceinline__ Fp fp_r2()
{
    constexpr uint32_t C[12] = {0x1c341746u, 0xf4df1f34u, 0x09d104f1u, 0x0a76e6a6u, 0x4c95b6d5u, 0x8de5476cu,
                                0x939d83c0u, 0x67eb88a9u, 0xb519952du, 0x9a793e85u, 0x92cae3aau, 0x11988fe5u};
    return fp_const(C);
}

// b = 4 of y^2 = x^3 + 4, Montgomery form
__device__ __forceinline__ Fp fp_b()
{
    constexpr uint32_t C[12] = {0x000cfff3u, 0xaa270000u, 0xfc34000au, 0x53cc0032u, 0x6b0a807fu, 0x478fe97au,
                                0xe6ba24d7u, 0xb1d37ebeu, 0xbf78ab2fu, 0x8ec9733bu, 0x3d83de7eu, 0x09d64551u};
    return fp_const(C);
}

// a^((p+1)/4): the square root of a when a is a square (p = 3 mod 4)
__device__ __forceinline__ Fp fp_sqrt_candidate(const Fp &a)
{
    constexpr uint32_t E[12] = {0xffffeaabu, 0xee7fbfffu, 0xac54ffffu, 0x07aaffffu, 0x3dac3d89u, 0xd9cc34a8u,
                                0x3ce144afu, 0xd91dd2e1u, 0x90d2eb35u, 0x92c6e9edu, 0x8e5ff9a6u, 0x0680447au};
    Fp r = Fp::one();
    for (int i = 378; i >= 0; --i) {
        r = sqr<FpCfg>(r);
        if ((E[i >> 5] >> (i & 31)) & 1u) r = r * a;
    }
    return r;
}

// 48 big-endian bytes (top three bits already cleared) -> canonical limbs;
// false when the value is not below p
__device__ __forceinline__ bool be_to_canonical(const uint8_t *b, Fp &x)
{
#pragma unroll
    for (int w = 0; w < 12; ++w) {
        const uint8_t *q = b + 44 - 4 * w;
        x.v[w] = (uint32_t)q[0] << 24 | (uint32_t)q[1] << 16 | (uint32_t)q[2] << 8 | (uint32_t)q[3];
    }
    for (int w = 11; w >= 0; --w)
        if (x.v[w] != FpCfg::p(w)) return x.v[w] < FpCfg::p(w);
    return false;
}

__device__ __forceinline__ bool canonical_gt(const Fp &a, const Fp &b)
{
    for (int w = 11; w >= 0; --w)
        if (a.v[w] != b.v[w]) return a.v[w] > b.v[w];
    return false;
}

// [r]P == O, by double-and-add over the bits of the group order
__device__ bool in_subgroup(const Affine &p)
{
    constexpr uint32_t Rr[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                                0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
    Xyzz acc = Xyzz::inf();
    for (int i = 254; i >= 0; --i) {
        acc = xyzz_dbl(acc);
        if ((Rr[i >> 5] >> (i & 31)) & 1u) acc = xyzz_add_affine(acc, p);
    }
    return acc.is_inf();
}

// one lane per point of g and g_lagrange (2n points); bad[0] counts the
// invalid ones, bad[1] holds the lowest invalid index (initialised to ~0)
__global__ void k_parse_g1(const uint8_t *in, size_t n2, uint32_t format, uint32_t *rows_g, uint32_t *rows_gl,
                           size_t n, uint32_t *bad)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n2) return;
    const bool compressed = format == kProcessed;
    const size_t pb = compressed ? 48 : 96;
    uint8_t b[96];
    for (size_t t = 0; t < pb; ++t) b[t] = in[i * pb + t];
    const uint32_t flag_c = b[0] >> 7, flag_inf = (b[0] >> 6) & 1u, flag_s = (b[0] >> 5) & 1u;
    b[0] &= 0x1f;
    bool ok = flag_c == (compressed ? 1u : 0u);
    Affine a;
    a.x = Fp::zero();
    a.y = Fp::zero();
    if (ok && flag_inf) {
        // the identity: no sign bit, every other bit zero
        uint32_t o = flag_s;
        for (size_t t = 0; t < pb; ++t) o |= b[t];
        ok = o == 0;
    } else if (ok && compressed) {
        Fp x;
        ok = be_to_canonical(b, x);
        if (ok) {
            a.x = x * fp_r2();
            const Fp rhs = sqr<FpCfg>(a.x) * a.x + fp_b();
            Fp y = fp_sqrt_candidate(rhs);
            ok = sqr<FpCfg>(y).eq(rhs);
            if (ok) {
                // the sign bit marks the lexicographically larger of y, -y
                const Fp ny = neg<FpCfg>(y);
                Fp one_raw = Fp::zero();
                one_raw.v[0] = 1;
                const bool larger = canonical_gt(y * one_raw, ny * one_raw);
                a.y = (larger == (flag_s != 0)) ? y : ny;
                ok = in_subgroup(a);
            }
        }
    } else if (ok) {
        Fp x, y;
        ok = flag_s == 0 && be_to_canonical(b, x) && be_to_canonical(b + 48, y);
        if (ok) {
            a.x = x * fp_r2();
            a.y = y * fp_r2();
            if (format == kRawBytes) {
                ok = sqr<FpCfg>(a.y).eq(sqr<FpCfg>(a.x) * a.x + fp_b());
                if (ok) ok = in_subgroup(a);
            }
        }
    }
    if (!ok) {
        atomicAdd(&bad[0], 1u);
        atomicMin(&bad[1], (uint32_t)i);
        a.x = Fp::zero();
        a.y = Fp::zero();
    }
    uint32_t *row = (i < n ? rows_g + i * 24 : rows_gl + (i - n) * 24);
    store_affine(row, a);
}

}  // namespace

struct Srs {
    uint32_t k = 0;
    std::unique_ptr<MsmContext> g, g_lagrange;
    std::vector<uint8_t> g2;   // G2 || [tau]G2 as read
};

std::unique_ptr<Srs> srs_load(const uint8_t *bytes, size_t len, uint32_t format)
{
    MH_CHECK(format <= kRawBytesUnchecked, (int)hipErrorInvalidValue, "unknown SRS serde format");
    MH_CHECK(len >= 4, (int)hipErrorInvalidValue, "SRS file too short");
    const uint32_t k = (uint32_t)bytes[0] | (uint32_t)bytes[1] << 8 | (uint32_t)bytes[2] << 16 |
                       (uint32_t)bytes[3] << 24;
    MH_CHECK(k <= 28, (int)hipErrorInvalidValue, "SRS degree above 2^28");
    const size_t n = size_t(1) << k;
    const size_t pb = format == kProcessed ? 48 : 96, g2b = format == kProcessed ? 96 : 192;
    const size_t need = 4 + 2 * n * pb + 2 * g2b;
    MH_CHECK(len >= need, (int)hipErrorInvalidValue, "SRS file shorter than its degree requires");
    hipStream_t st = cur_stream();
    auto srs = std::make_unique<Srs>();
    srs->k = k;
    srs->g = make_context(n);
    srs->g_lagrange = make_context(n);
    srs->g2.assign(bytes + 4 + 2 * n * pb, bytes + need);
    DevBuf<uint8_t> raw(2 * n * pb);
    DevBuf<uint32_t> bad(2);
    const uint32_t init[2] = {0u, ~0u};
    MH_HIP(hipMemcpyAsync(bad.get(), init, 8, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(raw.get(), bytes + 4, 2 * n * pb, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_parse_g1, dim3((unsigned)ceil_div(2 * n, 128)), dim3(128), 0, st, raw.get(), 2 * n, format,
                       srs->g->points.get(), srs->g_lagrange->points.get(), n, bad.get());
    MH_HIP(hipGetLastError());
    uint32_t res[2];
    MH_HIP(hipMemcpyAsync(res, bad.get(), 8, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    if (res[0]) {
        const bool lag = res[1] >= n;
        throw hip_error(-1, std::string("invalid point encoding in SRS file (") + std::to_string(res[0]) +
                                " invalid; first: " + (lag ? "g_lagrange[" : "g[") +
                                std::to_string(lag ? res[1] - n : res[1]) + "])");
    }
    build_precomp(*srs->g, st);
    build_precomp(*srs->g_lagrange, st);
    MH_HIP(hipStreamSynchronize(st));
    return srs;
}

}  // namespace mh

using namespace mh;

extern "C" RustError midnight_hip_srs_load(void **out, const uint8_t *bytes, size_t len, uint32_t format)
{
    if (!out || !bytes) return make_error((int)hipErrorInvalidValue, "invalid SRS load arguments");
    *out = nullptr;
    MH_TRY
    *out = srs_load(bytes, len, format).release();
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_hip_srs_load_file(void **out, const char *path, uint32_t format)
{
    if (!out || !path) return make_error((int)hipErrorInvalidValue, "invalid SRS load arguments");
    *out = nullptr;
    MH_TRY
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    MH_CHECK(f.good(), (int)hipErrorFileNotFound, std::string("cannot open SRS file ") + path);
    const std::streamoff size = f.tellg();
    MH_CHECK(size >= 0, (int)hipErrorFileNotFound, std::string("cannot read SRS file ") + path);
    std::vector<uint8_t> data((size_t)size);
    f.seekg(0);
    f.read(reinterpret_cast<char *>(data.data()), size);
    MH_CHECK(f.good() || f.eof(), (int)hipErrorFileNotFound, std::string("cannot read SRS file ") + path);
    *out = srs_load(data.data(), data.size(), format).release();
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_hip_srs_info(const void *srs, uint32_t *k, void **g_context, void **g_lagrange_context)
{
    if (!srs) return make_error((int)hipErrorInvalidValue, "SRS is null");
    auto *s = const_cast<Srs *>(static_cast<const Srs *>(srs));
    if (k) *k = s->k;
    if (g_context) *g_context = s->g.get();
    if (g_lagrange_context) *g_lagrange_context = s->g_lagrange.get();
    return ok();
}

extern "C" RustError midnight_hip_srs_g2(const void *srs, uint8_t *out, size_t capacity, size_t *len)
{
    if (!srs || !len) return make_error((int)hipErrorInvalidValue, "invalid SRS G2 arguments");
    const auto *s = static_cast<const Srs *>(srs);
    *len = s->g2.size();
    if (!out || capacity < s->g2.size()) return make_error((int)hipErrorInvalidValue, "G2 buffer too small");
    std::memcpy(out, s->g2.data(), s->g2.size());
    return ok();
}

extern "C" void midnight_hip_srs_free(void *srs)
{
    auto *s = static_cast<Srs *>(srs);
    if (!s) return;
    try {
        (void)hipStreamSynchronize(cur_stream());
    } catch (...) {
    }
    delete s;
}
