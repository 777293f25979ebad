// Poseidon hash chains and their circuit witness, and the scatter of that
// witness into the eight-column advice layout (bls12_381_g1_msm.cu:2018-2072,
// :2089-2216, :2820-2876, :4121-4300).
//
// Width-3 Poseidon with 4 + 4 full rounds around 60 partial rounds, S-box
// x^5, caller-supplied MDS (row-major 3 x 3) and round constants (68 x 3).
// A chain is strictly sequential -- every hash starts from the previous
// output -- so one wavefront runs it and what counts is the latency of a
// round.  Every lane keeps its own copy of one state element state[c]:
//   lanes 0..8   (l = k / 3, c = k % 3) own the MDS product mds[l][c] * s_c,
//   lanes 9..11  (c = k - 9) square state[c] for the product lanes and form
//                the S-box cubes the witness records.
// With the product taken as ((m x) x^2) x^2 = m x^5, a full round is three
// multiplication steps that every lane runs at once -- (m x | x^2),
// (m x^3 | x^3), (m x^5) -- with one LDS exchange of x^2 between steps one
// and two, where forming x^5 first and multiplying by m after is four; five
// needs three steps in any order (2^2 < 5).  A second LDS exchange sums the
// products: state'[c] = rc[r + 1][c] + sum_j prod[c][j].  Partial rounds
// take the same steps with only column 2 nonlinear.  The 204 round
// constants sit in LDS.
//
// Products are the radix-2^29 ones of fr29.hpp (~0.6x the instructions of
// 32-bit CIOS; issue count is what bounds a single wave).  Their Montgomery
// radix is 2^261, so the chain keeps every value as v 2^261 mod r: the
// parameters and inputs are scaled by 2^5 on the host, products of two such
// values stay in that form, sums do not depend on the radix, and one
// data-parallel pass after the chain brings outputs and witness back to
// the usual 2^256 form (a product by 2^256 mod r).
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "devvec.hpp"
#include "fields.hpp"
#include "fr29.hpp"
#include "host_fr.hpp"
#include "profile.hpp"
#include "runtime.hpp"

namespace mh {

namespace {

constexpr int kWidth = 3;
constexpr int kRounds = 68;
constexpr size_t kPermWitness = 128;
constexpr size_t kHashWitness = 2 * kPermWitness;
constexpr size_t kAdviceColumns = 8;
constexpr size_t kRowsPerHash = 41;
constexpr size_t kMetadataLen = 11;
constexpr size_t kLayoutLen = 128;

__device__ __forceinline__ bool is_full_round(int r) { return r < 4 || r >= 64; }

// One permutation.  `x` is this lane's copy of state[c]; `w` (nullable)
// receives the permutation's 128 witness values (in the 2^261 form): the
// cubes from the squaring lanes, the post-MDS states from the product lanes
// of row 0.
__device__ __forceinline__ void permute(Fr29 &x, const Fr29 &m, int l, int c, int k, const Fr29 *s_rc,
                                        Fr29 *s_sq, Fr29 (*s_prod)[9], Fr *w)
{
    x = fr29_add(x, s_rc[c]);
    size_t off = 0;
    const bool product_lane = k < 9;
    const bool writer = w != nullptr && product_lane && l == 0;
    for (int r = 0; r < kRounds; ++r) {
        const bool full = is_full_round(r);
        const bool nonlinear = full || c == 2;
        // step 1: m x on the product lanes, x^2 on the squaring lanes
        Fr29 y = fr29_mul(product_lane ? m : x, x);
        if (!product_lane) s_sq[c] = y;
        __syncthreads();
        if (nonlinear) {
            // step 2: m x^3 | x^3;  step 3: m x^5
            const Fr29 sq = product_lane ? s_sq[c] : y;
            y = fr29_mul(y, product_lane ? sq : x);
            if (!product_lane && full && w != nullptr) store_fr(w + off + c, fr29_to32(y));
            if (product_lane) y = fr29_mul(y, sq);
        }
        Fr29 *prod = s_prod[r & 1];
        if (product_lane) prod[k] = y;
        __syncthreads();
        Fr29 nx = s_rc[(r + 1) * kWidth + c];
        nx = fr29_add(nx, prod[c * kWidth + 0]);
        nx = fr29_add(nx, prod[c * kWidth + 1]);
        nx = fr29_add(nx, prod[c * kWidth + 2]);
        if (writer) {
            if (full) {
                store_fr(w + off + 3 + c, fr29_to32(nx));
            } else {
                int in_batch = (r - 4) % 6;
                if (in_batch < 5) {
                    if (c == 2) store_fr(w + off + in_batch, fr29_to32(nx));
                } else {
                    store_fr(w + off + 5 + c, fr29_to32(nx));
                }
            }
        }
        off += full ? 6 : ((r - 4) % 6 == 5 ? 8 : 0);
        x = nx;
    }
}

// One block of 64 threads (a single wavefront); lanes 12..63 only help load
// the round constants.  Parameters and inputs arrive in the 2^261 form.
__global__ __launch_bounds__(64) void k_poseidon_chain(Fr *out, Fr *witness, Fr initial, Fr tail0, Fr tail1,
                                                       Fr input_len, const Fr *mds, const Fr *rc, size_t len)
{
    // round constants with three zeros after the last round (its MDS output
    // gets no constant)
    __shared__ Fr29 s_rc[(kRounds + 1) * kWidth];
    __shared__ Fr29 s_prod[2][9];
    __shared__ Fr29 s_sq[kWidth];
    const int k = threadIdx.x;
    for (int i = k; i < (kRounds + 1) * kWidth; i += blockDim.x)
        s_rc[i] = fr29_from32(i < kRounds * kWidth ? load_fr(rc + i) : Fr::zero());
    __syncthreads();
    if (k >= 12) return;
    const int l = k < 9 ? k / kWidth : 0, c = k < 9 ? k % kWidth : k - 9;
    const Fr29 m = fr29_from32(k < 9 ? load_fr(mds + l * kWidth + c) : Fr::zero());
    const Fr29 t0 = fr29_from32(tail0), t1 = fr29_from32(tail1), il = fr29_from32(input_len);
    Fr29 previous = fr29_from32(initial);
    for (size_t index = 0; index < len; ++index) {
        Fr29 x = c == 2 ? il : (c == 0 ? previous : t0);
        Fr *w = witness ? witness + index * kHashWitness : nullptr;
        permute(x, m, l, c, k, s_rc, s_sq, s_prod, w);
        if (c == 0) x = fr29_add(x, t1);
        permute(x, m, l, c, k, s_rc, s_sq, s_prod, w ? w + kPermWitness : nullptr);
        previous = x;  // meaningful on the c == 0 lanes
        if (k == 0 && out) store_fr(out + index, fr29_to32(previous));
    }
}

// v 2^261 -> v 2^256 in place: the radix-2^29 product by 2^256 mod r
__global__ void k_from261(Fr *v, size_t n, Fr r256)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        store_fr(v + i, fr_mul_scaled(load_fr(v + i), r256));
}

// Witness scatter: value idx of hash h = idx / 256 goes to column
// layout >> 8, row 2 + 41 h + 21 p + (layout & 0xff) of the permutation p it
// belongs to.
__global__ void k_advice_scatter(Fr *out, size_t domain_len, const Fr *witness, size_t hash_count,
                                 const uint16_t *layout)
{
    const size_t total = hash_count * kHashWitness;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const size_t h = idx / kHashWitness, local = idx % kHashWitness, p = local / kPermWitness;
        const uint16_t packed = layout[local % kPermWitness];
        const size_t col = packed >> 8, row = packed & 0xff;
        store_fr(out + col * domain_len + 2 + kRowsPerHash * h + 21 * p + row, load_fr(witness + idx));
    }
}

// Rows fixed by the chain's inputs: each permutation's input state and the
// state plus the first round constants, the row adding tail1 between the
// two permutations, and the caller's prefix rows 0 and 1.  Metadata =
// {prefix[3], binding, initial, tail0, tail1, input_len, rc0[3]}.
__global__ void k_advice_inputs(Fr *out, size_t domain_len, const Fr *witness, size_t hash_count, const Fr *meta)
{
    for (size_t h = blockIdx.x * (size_t)blockDim.x + threadIdx.x; h < hash_count;
         h += (size_t)gridDim.x * blockDim.x) {
        const size_t r0 = 2 + kRowsPerHash * h;
        const Fr first[3] = {h == 0 ? load_fr(meta + 4) : load_fr(witness + (h - 1) * kHashWitness + 253),
                             load_fr(meta + 5), load_fr(meta + 7)};
        for (int c = 0; c < 3; ++c) {
            store_fr(out + c * domain_len + r0, first[c]);
            store_fr(out + c * domain_len + r0 + 1, first[c] + load_fr(meta + 8 + c));
        }
        const Fr *fo = witness + h * kHashWitness + 125;  // first permutation's output
        const Fr o0 = load_fr(fo), t1 = load_fr(meta + 6);
        const Fr second[3] = {o0 + t1, load_fr(fo + 1), load_fr(fo + 2)};
        store_fr(out + r0 + 20, second[0]);
        store_fr(out + domain_len + r0 + 20, o0);
        store_fr(out + 2 * domain_len + r0 + 20, t1);
        for (int c = 0; c < 3; ++c) {
            store_fr(out + c * domain_len + r0 + 21, second[c]);
            store_fr(out + c * domain_len + r0 + 22, second[c] + load_fr(meta + 8 + c));
        }
        if (h == 0) {
            store_fr(out, load_fr(meta + 0));
            store_fr(out + domain_len, load_fr(meta + 1));
            store_fr(out + 2 * domain_len, load_fr(meta + 2));
            store_fr(out + 1, load_fr(meta + 3));
        }
    }
}

bool chain_args_ok(const RawFr *out, const RawFr *mds, size_t mds_len, const RawFr *rc, size_t rc_len)
{
    return out != nullptr && mds != nullptr && mds_len == (size_t)kWidth * kWidth && rc != nullptr &&
           rc_len == (size_t)kRounds * kWidth;
}

// Runs the chain; `with_witness` selects the witness stream (len * 256
// values) instead of the outputs (len values) as what is copied to `out`.
// v (usual 2^256 form) -> v 2^261 mod r, which is the usual form of 32 v
Fr scaled261(const RawFr &v)
{
    const HostFr h = to_hfr(v) * HostFr::from_u64(32);
    Fr r;
    std::memcpy(r.v, h.l, 32);
    return r;
}

void run_chain(RawFr *out, size_t len, const RawFr &initial, const RawFr &tail0, const RawFr &tail1,
               const RawFr &input_len, const RawFr *mds, const RawFr *rc, bool with_witness)
{
    hipStream_t st = cur_stream();
    const size_t n_out = with_witness ? len * kHashWitness : len;
    constexpr size_t kParams = kWidth * kWidth + kRounds * kWidth;
    std::vector<Fr> params(kParams);
    for (size_t i = 0; i < kParams; ++i) params[i] = scaled261(i < 9 ? mds[i] : rc[i - 9]);
    DevBuf<Fr> d_out(n_out), d_params(kParams);
    MH_HIP(hipMemcpyAsync(d_params.get(), params.data(), kParams * 32, hipMemcpyHostToDevice, st));
    ProfileScope prof;
    prof.begin("poseidon_chain", st);
    hipLaunchKernelGGL(k_poseidon_chain, dim3(1), dim3(64), 0, st, with_witness ? nullptr : d_out.get(),
                       with_witness ? d_out.get() : nullptr, scaled261(initial), scaled261(tail0),
                       scaled261(tail1), scaled261(input_len), (const Fr *)d_params.get(),
                       (const Fr *)d_params.get() + kWidth * kWidth, len);
    MH_HIP(hipGetLastError());
    prof.end(st);
    Fr r256;
    const HostFr one = HostFr::one();
    std::memcpy(r256.v, one.l, 32);
    hipLaunchKernelGGL(k_from261, dim3((unsigned)std::min<size_t>(ceil_div(n_out, 256), 8192)), dim3(256), 0, st,
                       d_out.get(), n_out, r256);
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(out, d_out.get(), n_out * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    prof.flush();
}

}  // namespace

}  // namespace mh

using namespace mh;

// bls12_381_g1_msm.cu:4121-4190
extern "C" RustError midnight_sppark_bls12_381_fr_poseidon_hash_chain(RawFr *out, size_t len, RawFr initial,
                                                                     RawFr tail0, RawFr tail1, RawFr input_len,
                                                                     const RawFr *mds, size_t mds_len,
                                                                     const RawFr *round_constants,
                                                                     size_t round_constants_len)
{
    if (len == 0) return ok();
    if (!chain_args_ok(out, mds, mds_len, round_constants, round_constants_len))
        return make_error((int)hipErrorInvalidValue, "invalid Poseidon hash-chain arguments");
    MH_TRY
    run_chain(out, len, initial, tail0, tail1, input_len, mds, round_constants, false);
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:4192-4300; out holds len * 256 values
extern "C" RustError midnight_sppark_bls12_381_fr_poseidon_hash_chain_witness(RawFr *out, size_t len, RawFr initial,
                                                                             RawFr tail0, RawFr tail1,
                                                                             RawFr input_len, const RawFr *mds,
                                                                             size_t mds_len,
                                                                             const RawFr *round_constants,
                                                                             size_t round_constants_len)
{
    if (len == 0) return ok();
    if (!chain_args_ok(out, mds, mds_len, round_constants, round_constants_len) || len > SIZE_MAX / 32 / kHashWitness)
        return make_error((int)hipErrorInvalidValue, "invalid Poseidon witness arguments");
    MH_TRY
    run_chain(out, len, initial, tail0, tail1, input_len, mds, round_constants, true);
    return ok();
    MH_CATCH
}

// bls12_381_g1_msm.cu:2820-2876.  Every layout entry must name a column
// below 8 and a row that stays inside the domain for the last hash; the
// reference leaves that to its caller (circuit.rs:838-846 checks the column
// and a row below 256).
extern "C" RustError midnight_sppark_bls12_381_fr_poseidon_advice_from_witness(
    void **out, const RawFr *witness, size_t hash_count, size_t domain_len, const RawFr *metadata,
    size_t metadata_len, const uint16_t *layout, size_t layout_len)
{
    if (out == nullptr || hash_count == 0 || domain_len == 0 || witness == nullptr || metadata == nullptr ||
        metadata_len != kMetadataLen || layout == nullptr || layout_len != kLayoutLen ||
        hash_count > SIZE_MAX / kHashWitness / 32 || domain_len > SIZE_MAX / kAdviceColumns / 32 ||
        hash_count > (SIZE_MAX - 2) / kRowsPerHash || 2 + kRowsPerHash * hash_count >= domain_len)
        return make_error((int)hipErrorInvalidValue, "invalid Poseidon advice arguments");
    const size_t last_start = 2 + kRowsPerHash * (hash_count - 1) + 21;
    for (size_t i = 0; i < kLayoutLen; ++i)
        if ((layout[i] >> 8) >= kAdviceColumns || last_start + (layout[i] & 0xff) >= domain_len)
            return make_error((int)hipErrorInvalidValue, "invalid Poseidon advice arguments");
    *out = nullptr;
    MH_TRY
    hipStream_t st = cur_stream();
    const size_t wlen = hash_count * kHashWitness;
    auto dv = std::make_unique<DeviceVec>();
    dv->lane = current_lane();
    dv->len = kAdviceColumns * domain_len;
    dv->values.alloc(dv->len);
    DevBuf<Fr> d_w(wlen), d_meta(kMetadataLen);
    DevBuf<uint16_t> d_layout(kLayoutLen);
    MH_HIP(hipMemsetAsync(dv->values.get(), 0, dv->len * 32, st));
    MH_HIP(hipMemcpyAsync(d_w.get(), witness, wlen * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(d_meta.get(), metadata, kMetadataLen * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(d_layout.get(), layout, kLayoutLen * 2, hipMemcpyHostToDevice, st));
    size_t grid = std::min<size_t>(ceil_div(wlen, 256), 65535u * 16u);
    hipLaunchKernelGGL(k_advice_scatter, dim3((unsigned)grid), dim3(256), 0, st, dv->values.get(), domain_len,
                       (const Fr *)d_w.get(), hash_count, (const uint16_t *)d_layout.get());
    MH_HIP(hipGetLastError());
    grid = std::min<size_t>(ceil_div(hash_count, 256), 65535u);
    hipLaunchKernelGGL(k_advice_inputs, dim3((unsigned)grid), dim3(256), 0, st, dv->values.get(), domain_len,
                       (const Fr *)d_w.get(), hash_count, (const Fr *)d_meta.get());
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(st));
    *out = dv.release();
    return ok();
    MH_CATCH
}
