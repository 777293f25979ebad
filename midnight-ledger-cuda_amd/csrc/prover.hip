// Proof pipeline: key generation and create_proof for PLONK circuits made of
// custom gates, a copy permutation, lookups, trashcans and (non-committed)
// instance columns, device-resident from the advice upload to the last
// opening commitment.  The benchmark circuit family has 5 advice and 3 fixed
// columns, gate q(ab + c - d) and 8 permuted columns.
//
// Stage order, transcript and blinding draws follow the reference prover
// (vendor/midnight-proofs-0.8.1/src/plonk/prover.rs:68-575 compute_trace,
// :769-1010 finalise_proof, :1237-1330 compute_instances, :1931-2080
// evaluations and queries; lookup/prover.rs:737-1072 commit_permuted_batch,
// :1075-1340 products, :1341-1572 evaluate / open, :1573-1690
// permute_expression_pair; trash/prover.rs; permutation/prover.rs:327-530,
// :545-712; vanishing/prover.rs:41-406; poly/kzg/mod.rs:1978-2570
// multi_open; transcript/{mod.rs, implementors.rs}):
//   vk repr -> instance values (length, values) -> advice (blinded rows
//   u..n from the RNG) -> commit -> theta -> lookup permuted input / table
//   (blinds from the RNG) -> commit -> beta, gamma -> permutation products,
//   lookup products (blinds from the RNG) -> commit -> trash challenge ->
//   trash polynomials -> commit -> random polynomial (ChaCha20 seeded from
//   the RNG) -> commit -> y -> numerator on the extended coset -> h ->
//   pieces (telescoping blinds) -> commit -> x -> evaluations -> multi-open
//   (x1, x2, f, x3, q evaluations, x4, pi).
// Every polynomial stays on the device; the host holds the transcript, the
// RNG and the few field elements the protocol reads back (evaluations).
//
// The verifying-key representative hashes the caller's description bytes
// where the reference hashes the Debug text of the pinned domain and
// constraint system (plonk/mod.rs:371-405): a caller that passes that text
// gets the reference's representative.  The vanishing argument's random
// polynomial is seeded per chunk of the reference prover's thread pool
// (vanishing/prover.rs:41-75); the thread count is a key setting
// (midnight_hip_plonk_set_rng_threads, default 1).  Commitments are computed
// from coefficients with the monomial SRS (commit_lagrange with the Lagrange
// SRS gives the same points).

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "devvec.hpp"
#include "fields.hpp"
#include "frops.hpp"
#include "host_fr.hpp"
#include "host_g1.hpp"
#include "hostcrypto.hpp"
#include "msm.hpp"
#include "msm_context.hpp"
#include "ntt.hpp"
#include "numerator.hpp"
#include "polyops.hpp"
#include "profile.hpp"
#include "rngdev.hpp"
#include "runtime.hpp"

namespace mh {

void g1_compress(const RawG1Projective &p, uint8_t out[48])
{
    HFp X, Y, Z;
    std::memcpy(X.l, p.x, 48);
    std::memcpy(Y.l, p.y, 48);
    std::memcpy(Z.l, p.z, 48);
    std::memset(out, 0, 48);
    if (Z.is_zero()) {
        out[0] = 0xc0;
        return;
    }
    const HFp zi = Z.inverse();
    const HFp zi2 = zi.sqr();
    const HFp x = X * zi2, y = Y * zi2 * zi;
    uint64_t xc[6], yc[6], nc[6];
    x.to_canonical(xc);
    y.to_canonical(yc);
    (HFp::zero() - y).to_canonical(nc);
    bool larger = false;   // y > -y as integers
    for (int i = 5; i >= 0; --i) {
        if (yc[i] != nc[i]) {
            larger = yc[i] > nc[i];
            break;
        }
    }
    for (int i = 0; i < 48; ++i) out[i] = (uint8_t)(xc[5 - i / 8] >> (8 * (7 - i % 8)));
    out[0] |= 0x80 | (larger ? 0x20 : 0);
}

namespace {

const uint8_t kTranscriptKey[] = "Domain separator for transcript";

// Fiat-Shamir transcript over Blake2b-512 (transcript/implementors.rs:13-33):
// absorb = 0x01 || bytes, squeeze = 0x00 then the digest of the running
// state (which continues), written values are appended to the proof.
class Transcript {
public:
    Transcript() : h_(64, kTranscriptKey, sizeof(kTranscriptKey) - 1, nullptr) {}
    void absorb(const uint8_t *b, size_t n)
    {
        const uint8_t p = 1;
        h_.update(&p, 1);
        h_.update(b, n);
    }
    void common_scalar(const HostFr &f)
    {
        uint8_t b[32];
        fr_to_bytes(f, b);
        absorb(b, 32);
    }
    void common_point(const RawG1Projective &p)
    {
        uint8_t b[48];
        g1_compress(p, b);
        absorb(b, 48);
    }
    void write_scalar(const HostFr &f)
    {
        uint8_t b[32];
        fr_to_bytes(f, b);
        absorb(b, 32);
        proof.insert(proof.end(), b, b + 32);
    }
    void write_point(const RawG1Projective &p)
    {
        uint8_t b[48];
        g1_compress(p, b);
        absorb(b, 48);
        proof.insert(proof.end(), b, b + 48);
    }
    HostFr squeeze()
    {
        const uint8_t p = 0;
        h_.update(&p, 1);
        const std::vector<uint8_t> d = h_.digest();
        return fr_from_uniform_bytes(d.data());
    }
    std::vector<uint8_t> proof;

private:
    Blake2b h_;
};

void check(const RustError &e)
{
    if (e.code != 0) {
        std::string m = e.message ? e.message : "backend error";
        free(e.message);
        throw hip_error(e.code, m);
    }
}

RawFr raw(const HostFr &h)
{
    RawFr r;
    std::memcpy(r.limbs, h.l, 32);
    return r;
}

HostFr hpow(const HostFr &a, uint64_t e)
{
    const uint64_t ee[4] = {e, 0, 0, 0};
    return a.pow(ee);
}

// ZETA (cube root of unity) and DELTA of the scalar field, Montgomery form
// (vendor/midnight-curves-0.3.1/src/bls12_381/fq.rs:492-500, :693-702)
const HostFr kZeta{{0x92d9090b093011d2ULL, 0xfc9cbd719d6aa073ULL, 0xc1f14ef0cd65a1a6ULL, 0x017f6d35e72fcdebULL}};
const HostFr kDelta{{0x70e310d3d146f96aULL, 0x4b64c08919e299e6ULL, 0x51e114186a8b970dULL, 0x6185d06627c067cbULL}};

std::unique_ptr<DeviceVec> new_vec(size_t len)
{
    auto v = std::make_unique<DeviceVec>();
    v->lane = current_lane();
    v->len = len;
    v->values.alloc(len);
    return v;
}

void ntt_batch(Fr *data, uint32_t log_n, size_t count, bool inverse, DevBuf<Fr> &scratch, hipStream_t st)
{
    scratch.ensure(ntt_scratch_elems(log_n, count));
    NttJob j;
    j.in = data;
    j.out = data;
    j.scratch = scratch.get();
    j.in_pstride = j.out_pstride = size_t(1) << log_n;
    j.log_n = log_n;
    j.count = (uint32_t)count;
    j.inverse = inverse;
    ntt_run(j, st);
}

// coefficients (len n each, count polys) -> values on the extended coset;
// `scaled` writes them in the scaled form v 2^261 mod r (times 32), the form
// the numerator computes in
void to_extended(Fr *dst, const Fr *src, size_t n, size_t count, uint32_t ext_log_n, DevBuf<Fr> &scratch,
                 hipStream_t st, bool scaled = false)
{
    scratch.ensure(ntt_scratch_elems(ext_log_n, count));
    NttJob j;
    j.in = src;
    j.out = dst;
    j.scratch = scratch.get();
    j.in_pstride = n;
    j.out_pstride = size_t(1) << ext_log_n;
    j.log_n = ext_log_n;
    j.count = (uint32_t)count;
    j.pre = 1;
    j.in_len = n;
    j.first = kZeta;
    j.second = kZeta.sqr();
    if (scaled) {
        j.post_mul = true;
        j.mul = HostFr::from_u64(32);
    }
    ntt_run(j, st);
}

}  // namespace

// A commitment stage of the proof pipeline captured once per key and lane in
// a HIP graph: the inverse NTT of the stage's polynomials in place (when the
// stage has one), then the MSM batch of those polynomials against the SRS.
// Every proof on the lane replays it: one launch instead of the NTT passes
// and the MSM's ~20 kernels, and no read-back inside the MSM.  The stage owns
// its MSM engine, so nothing the graph refers to is reallocated by other
// calls.
struct StageGraph {
    MsmEngine engine;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    StageGraph() = default;
    StageGraph(const StageGraph &) = delete;
    StageGraph &operator=(const StageGraph &) = delete;
    ~StageGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// Per-lane state of the proofs of one key: the polynomial arena and the
// [f | work] pair at fixed addresses (which the captured stages refer to),
// the NTT scratch of the captured transforms, and the stages themselves.
struct LaneProof {
    std::unique_ptr<DeviceVec> polys, fv;
    DevBuf<Fr> ntt_scratch;
    StageGraph advice, lookups, products, trash, random, pieces, opening;
};

struct PlonkKey {
    uint32_t lane = 0;
    uint32_t k = 0, ext_k = 0, degree = 0;
    size_t n = 0, ext_n = 0, bf = 0, u = 0;
    size_t fixed_cols = 0, advice_cols = 0, perm_cols = 0, chunk = 0, sets = 0, pieces = 0;
    size_t instance_cols = 0, committed = 0, lookups = 0, trashes = 0;
    std::vector<std::pair<uint32_t, int32_t>> committed_queries;   // instance queries of committed columns
    void *srs = nullptr;                                   // MSM context (not owned), >= n points
    std::vector<std::pair<uint32_t, int32_t>> advice_queries, fixed_queries;
    std::vector<RawColumnRef> perm_columns;
    std::unique_ptr<DeviceVec> fixed_coeff, sigma_coeff;  // coefficient form
    std::vector<RawG1Projective> fixed_com, sigma_com;
    HostFr omega, omega_inv, ext_omega, repr;
    std::vector<HostFr> t_inv, t_inv_scaled, delta_powers;
    void *numctx = nullptr;
    uint32_t rng_threads = 1;   // thread count of the reference prover whose random draws are reproduced
    bool graphs = true;         // commitment stages replayed from HIP graphs
    DevBuf<Fr> scratch_by_lane[kLaneCount];   // NTT scratch of proofs running on each lane
    DevBuf<Fr> &scratch() { return scratch_by_lane[current_lane()]; }
    std::unique_ptr<LaneProof> lane_proofs[kLaneCount];
    ~PlonkKey()
    {
        for (auto &lp : lane_proofs) lp.reset();
        if (numctx) midnight_sppark_bls12_381_numerator_context_free(numctx);
    }
};

namespace {

void commit(std::vector<RawG1Projective> &out, PlonkKey &pk, const DeviceVec &v, size_t offset_polys, size_t count)
{
    out.resize(count);
    check(midnight_sppark_bls12_381_g1_msm_context_device_batch(out.data(), pk.srs, &v, offset_polys * pk.n, pk.n,
                                                                count));
}

// One captured commitment stage: `count` polynomials of n coefficients at
// `base` (transformed from Lagrange values first when `intt`), committed into
// out[0..count).  Captured on the first call, replayed after.
void stage_commit(StageGraph &g, PlonkKey &pk, LaneProof &lp, Fr *base, size_t count, bool intt,
                  std::vector<RawG1Projective> &out, hipStream_t st)
{
    const size_t n = pk.n;
    out.resize(count);
    if (!g.exec) {
        MsmContext &ctx = *static_cast<MsmContext *>(pk.srs);
        std::vector<MsmJob> jobs(count);
        for (size_t i = 0; i < count; ++i) jobs[i] = MsmJob{base + i * n, 0, n};
        // the same table / plain-window choice as the context's own batches
        g.engine.chunk_len = ctx.chunk_len;
        PrecompTable pc;
        if (context_table(ctx, n, n, pc)) {
            g.engine.prepare(ctx.points.get(), ctx.npoints, jobs, 0, st, &pc);
        } else {
            g.engine.prepare(ctx.points.get(), ctx.npoints, jobs, ctx.window_bits, st);
        }
        MH_CHECK(!intt || lp.ntt_scratch.len >= ntt_scratch_elems(pk.k, count), -1,
                 "captured transform larger than the lane's scratch");
        MH_HIP(hipStreamSynchronize(st));
        MH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            if (intt) ntt_batch(base, pk.k, count, true, lp.ntt_scratch, st);
            g.engine.enqueue(st, nullptr);
        } catch (...) {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(st, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            throw;
        }
        MH_HIP(hipStreamEndCapture(st, &g.graph));
        MH_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    }
    ProfileScope prof;
    prof.begin("commit_stage", st);
    MH_HIP(hipGraphLaunch(g.exec, st));
    prof.end(st);
    MH_HIP(hipStreamSynchronize(st));
    if (profiling()) prof.flush();
    g.engine.finish(out.data());
}

HostFr rotate(const PlonkKey &pk, const HostFr &x, int32_t rot)
{
    HostFr r = x;
    const HostFr w = rot >= 0 ? pk.omega : pk.omega_inv;
    for (int32_t i = 0; i < (rot >= 0 ? rot : -rot); ++i) r = r * w;
    return r;
}

}  // namespace

// --------------------------------------------------------------------- keygen

std::unique_ptr<PlonkKey> plonk_keygen(uint32_t k, void *srs, size_t srs_points, uint32_t degree, const RawFr *fixed,
                                       size_t fixed_cols, size_t advice_cols,
                                       const std::vector<std::pair<uint32_t, int32_t>> &advice_queries,
                                       const std::vector<std::pair<uint32_t, int32_t>> &fixed_queries,
                                       const RawColumnRef *perm_columns, size_t perm_cols, const uint32_t *mapping,
                                       const RawGraph &gate, const std::string &description,
                                       const RawPlonkArguments &args)
{
    hipStream_t st = cur_stream();
    auto pk = std::make_unique<PlonkKey>();
    pk->lane = current_lane();
    pk->k = k;
    pk->n = size_t(1) << k;
    pk->degree = degree;
    const size_t n = pk->n;
    MH_CHECK(srs_points >= n, (int)hipErrorInvalidValue, "SRS is smaller than the circuit");
    MH_CHECK(degree >= 3 && perm_cols > 0, (int)hipErrorInvalidValue, "circuit degree below 3 or no permutation");
    pk->srs = srs;
    // EvaluationDomain::new(degree, k) (poly/domain.rs:54-159)
    uint32_t ek = k;
    while ((size_t(1) << ek) < n * (degree - 1)) ++ek;
    pk->ext_k = ek;
    pk->ext_n = size_t(1) << ek;
    pk->pieces = degree - 1;
    // the transforms of both sizes find their twiddle tables built, also
    // the ones captured in the commitment stages' graphs
    ntt_warm_tables(k, st);
    ntt_warm_tables(ek, st);
    pk->omega = HostFr::omega(k);
    pk->omega_inv = pk->omega.inverse();
    pk->ext_omega = HostFr::omega(ek);
    {
        const HostFr orig = hpow(kZeta, n), step = hpow(pk->ext_omega, n);
        HostFr cur = orig;
        for (size_t i = 0; i < (size_t(1) << (ek - k)); ++i) {
            pk->t_inv.push_back((cur - HostFr::one()).inverse());
            cur = cur * step;
        }
        // for a numerator written in the scaled form (32 times its value)
        const HostFr inv32 = HostFr::from_u64(32).inverse();
        for (const HostFr &t : pk->t_inv) pk->t_inv_scaled.push_back(t * inv32);
    }
    // blinding factors (plonk/circuit.rs:2502-2530): max(3, queries per advice column) + 1 + 1
    size_t maxq = 1;
    for (size_t c = 0; c < advice_cols; ++c) {
        size_t q = 0;
        for (auto &aq : advice_queries) q += aq.first == c;
        maxq = std::max(maxq, q);
    }
    pk->bf = std::max<size_t>(3, maxq) + 2;
    MH_CHECK(n > pk->bf + 1, (int)hipErrorInvalidValue, "circuit too small for its blinding rows");
    pk->u = n - pk->bf - 1;
    pk->fixed_cols = fixed_cols;
    pk->advice_cols = advice_cols;
    pk->perm_cols = perm_cols;
    pk->instance_cols = args.instance_cols;
    pk->committed = args.committed_instances;
    for (size_t i = 0; i < args.instance_query_count; ++i)
        if (args.instance_queries[2 * i] < args.committed_instances)
            pk->committed_queries.emplace_back(args.instance_queries[2 * i], (int32_t)args.instance_queries[2 * i + 1]);
    pk->lookups = args.lookup_count;
    pk->trashes = args.trash_count;
    for (size_t j = 0; j < perm_cols; ++j)
        MH_CHECK(perm_columns[j].kind <= 2 &&
                     perm_columns[j].index < (perm_columns[j].kind == 0   ? fixed_cols
                                              : perm_columns[j].kind == 1 ? advice_cols
                                                                          : args.instance_cols),
                 (int)hipErrorInvalidValue, "permuted column out of range");
    pk->chunk = degree - 2;
    pk->sets = (perm_cols + pk->chunk - 1) / pk->chunk;
    pk->advice_queries = advice_queries;
    pk->fixed_queries = fixed_queries;
    pk->perm_columns.assign(perm_columns, perm_columns + perm_cols);
    pk->delta_powers.resize(perm_cols);
    HostFr d = HostFr::one();
    for (size_t j = 0; j < perm_cols; ++j, d = d * kDelta) pk->delta_powers[j] = d;

    // sigma Lagrange values delta^j' omega^i' (permutation/keygen.rs)
    std::vector<HostFr> wpow(n);
    wpow[0] = HostFr::one();
    for (size_t i = 1; i < n; ++i) wpow[i] = wpow[i - 1] * pk->omega;
    std::vector<HostFr> sigma(perm_cols * n);
    for (size_t j = 0; j < perm_cols; ++j)
        for (size_t i = 0; i < n; ++i) {
            const uint32_t cj = mapping[2 * (j * n + i)], ri = mapping[2 * (j * n + i) + 1];
            MH_CHECK(cj < perm_cols && ri < n, (int)hipErrorInvalidValue, "permutation mapping out of range");
            sigma[j * n + i] = pk->delta_powers[cj] * wpow[ri];
        }
    // l0, l_last, l_active on the base domain
    std::vector<HostFr> lag(3 * n, HostFr::zero());
    lag[0] = HostFr::one();
    lag[n + pk->u] = HostFr::one();
    for (size_t i = 0; i < pk->u; ++i) lag[2 * n + i] = HostFr::one();

    // Lagrange -> coefficients -> commitments; extended cosets
    DeviceVec fixed_lag, sigma_lag;
    fixed_lag.len = fixed_cols * n;
    fixed_lag.values.alloc(std::max<size_t>(fixed_lag.len, 1));
    sigma_lag.len = perm_cols * n;
    sigma_lag.values.alloc(sigma_lag.len);
    if (fixed_cols) MH_HIP(hipMemcpyAsync(fixed_lag.values.get(), fixed, fixed_cols * n * 32, hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(sigma_lag.values.get(), sigma.data(), perm_cols * n * 32, hipMemcpyHostToDevice, st));
    pk->fixed_coeff = new_vec(std::max<size_t>(fixed_cols, 1) * n);
    pk->sigma_coeff = new_vec(perm_cols * n);
    pk->fixed_coeff->len = fixed_cols * n;
    if (fixed_cols)
        MH_HIP(hipMemcpyAsync(pk->fixed_coeff->values.get(), fixed_lag.values.get(), fixed_cols * n * 32,
                              hipMemcpyDeviceToDevice, st));
    MH_HIP(hipMemcpyAsync(pk->sigma_coeff->values.get(), sigma_lag.values.get(), perm_cols * n * 32,
                          hipMemcpyDeviceToDevice, st));
    if (fixed_cols) ntt_batch(pk->fixed_coeff->values.get(), k, fixed_cols, true, pk->scratch(), st);
    ntt_batch(pk->sigma_coeff->values.get(), k, perm_cols, true, pk->scratch(), st);
    DevBuf<Fr> lag_d(3 * n);
    MH_HIP(hipMemcpyAsync(lag_d.get(), lag.data(), 3 * n * 32, hipMemcpyHostToDevice, st));
    ntt_batch(lag_d.get(), k, 3, true, pk->scratch(), st);
    MH_HIP(hipStreamSynchronize(st));
    if (fixed_cols) commit(pk->fixed_com, *pk, *pk->fixed_coeff, 0, fixed_cols);
    commit(pk->sigma_com, *pk, *pk->sigma_coeff, 0, perm_cols);

    const size_t N = pk->ext_n;
    DevBuf<Fr> fixed_ext(std::max<size_t>(fixed_cols, 1) * N), sigma_ext(perm_cols * N), lag_ext(3 * N);
    if (fixed_cols) to_extended(fixed_ext.get(), pk->fixed_coeff->values.get(), n, fixed_cols, ek, pk->scratch(), st);
    to_extended(sigma_ext.get(), pk->sigma_coeff->values.get(), n, perm_cols, ek, pk->scratch(), st);
    to_extended(lag_ext.get(), lag_d.get(), n, 3, ek, pk->scratch(), st);
    MH_HIP(hipStreamSynchronize(st));
    // the numerator context keeps its own copies; host staging of the cosets
    std::vector<RawFr> h_fixed_ext(fixed_cols * N), h_sigma_ext(perm_cols * N), h_lag_ext(3 * N),
        h_fixed_lag(fixed_cols * n);
    if (fixed_cols) {
        MH_HIP(hipMemcpyAsync(h_fixed_ext.data(), fixed_ext.get(), fixed_cols * N * 32, hipMemcpyDeviceToHost, st));
        std::memcpy(h_fixed_lag.data(), fixed, fixed_cols * n * 32);
    }
    MH_HIP(hipMemcpyAsync(h_sigma_ext.data(), sigma_ext.get(), perm_cols * N * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipMemcpyAsync(h_lag_ext.data(), lag_ext.get(), 3 * N * 32, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    check(midnight_sppark_bls12_381_numerator_context_create(
        &pk->numctx, N, n, fixed_cols ? h_fixed_ext.data() : nullptr, fixed_cols ? h_fixed_lag.data() : nullptr,
        fixed_cols, h_lag_ext.data(), h_lag_ext.data() + N, h_lag_ext.data() + 2 * N, perm_columns, perm_cols,
        h_sigma_ext.data(), reinterpret_cast<const RawFr *>(sigma.data()), perm_cols, args.trash_selectors,
        args.trash_count, gate, args.lookup_graphs, args.lookup_input_graphs, args.lookup_table_graphs,
        args.lookup_count, args.trash_graphs, raw(pk->ext_omega), raw(pk->omega), raw(kZeta), raw(kDelta),
        (int32_t)(1u << (ek - k)), -(int32_t)(pk->bf + 1), (uint32_t)pk->chunk));

    // verifying-key representative (plonk/mod.rs:371-405): Blake2b-512
    // personalised "Halo2-Verify-Key" over version, k, the commitments
    // (96-byte affine Montgomery limbs) and a description of the circuit
    const uint8_t personal[16] = {'H', 'a', 'l', 'o', '2', '-', 'V', 'e', 'r', 'i', 'f', 'y', '-', 'K', 'e', 'y'};
    Blake2b hv(64, nullptr, 0, personal);
    const uint8_t head[2] = {0x03, (uint8_t)k};
    hv.update(head, 2);
    auto put_points = [&](const std::vector<RawG1Projective> &pts) {
        const uint32_t cnt = (uint32_t)pts.size();
        uint8_t c[4] = {(uint8_t)cnt, (uint8_t)(cnt >> 8), (uint8_t)(cnt >> 16), (uint8_t)(cnt >> 24)};
        hv.update(c, 4);
        for (const auto &p : pts) {
            HFp X, Y, Z;
            std::memcpy(X.l, p.x, 48);
            std::memcpy(Y.l, p.y, 48);
            std::memcpy(Z.l, p.z, 48);
            uint8_t b[96] = {0};
            if (!Z.is_zero()) {
                const HFp zi = Z.inverse(), zi2 = zi.sqr();
                const HFp x = X * zi2, y = Y * zi2 * zi;
                std::memcpy(b, x.l, 48);
                std::memcpy(b + 48, y.l, 48);
            }
            hv.update(b, 96);
        }
    };
    put_points(pk->fixed_com);
    put_points(pk->sigma_com);
    hv.update(reinterpret_cast<const uint8_t *>(description.data()), description.size());
    pk->repr = fr_from_uniform_bytes(hv.digest().data());
    return pk;
}

// HBM a proof of this key holds beyond the key itself: the extended
// advice and product cosets with the numerator and quotient, the
// coefficient-form polynomials and opening temporaries, and MSM scratch.
size_t plonk_footprint(const PlonkKey &pk)
{
    const size_t polys = pk.advice_cols + pk.sets + 1 + pk.pieces + 3 * pk.lookups + pk.trashes + pk.instance_cols;
    const size_t ext_cols = pk.advice_cols + pk.instance_cols + pk.sets + 3 * pk.lookups + pk.trashes + 2;
    return 32 * (pk.ext_n * ext_cols + pk.n * (2 * polys + 12 + 4 * pk.lookups + pk.trashes)) + 256 * pk.n;
}

size_t plonk_advice_len(const PlonkKey &pk) { return pk.advice_cols * pk.n; }
size_t plonk_instance_cols(const PlonkKey &pk) { return pk.instance_cols; }

// Length of every proof of this key: the commitments and evaluations that
// plonk_prove writes, with the multi-open's point sets formed as it forms
// them (queries at distinct rotations of x are distinct points).
size_t plonk_proof_size(const PlonkKey &pk)
{
    const size_t A = pk.advice_cols, S = pk.sets, L = pk.lookups, TR = pk.trashes;
    const int32_t last = -(int32_t)(pk.bf + 1);
    // (polynomial id, rotation) in the prover's query order; ids: advice,
    // permutation sets, lookup products, permuted inputs, permuted tables,
    // trash polynomials, fixed columns, sigmas, h, the random polynomial
    const long z0 = (long)A, lz0 = z0 + (long)S, li0 = lz0 + (long)L, lt0 = li0 + (long)L, tr0 = lt0 + (long)L;
    const long fixed0 = tr0 + (long)TR, sigma0 = fixed0 + (long)pk.fixed_cols, hid = sigma0 + (long)pk.perm_cols;
    const long inst0 = hid + 2;
    std::vector<std::pair<long, int32_t>> q;
    for (auto &a : pk.advice_queries) q.emplace_back((long)a.first, a.second);
    for (auto &c : pk.committed_queries) q.emplace_back(inst0 + (long)c.first, c.second);
    for (size_t s = 0; s < S; ++s) {
        q.emplace_back(z0 + (long)s, 0);
        q.emplace_back(z0 + (long)s, 1);
    }
    for (size_t s = S - 1; s-- > 0;) q.emplace_back(z0 + (long)s, last);
    for (size_t l = 0; l < L; ++l) {
        q.emplace_back(lz0 + (long)l, 0);
        q.emplace_back(li0 + (long)l, 0);
        q.emplace_back(lt0 + (long)l, 0);
        q.emplace_back(li0 + (long)l, -1);
        q.emplace_back(lz0 + (long)l, 1);
    }
    for (size_t t = 0; t < TR; ++t) q.emplace_back(tr0 + (long)t, 0);
    for (auto &f : pk.fixed_queries) q.emplace_back(fixed0 + f.first, f.second);
    for (size_t j = 0; j < pk.perm_cols; ++j) q.emplace_back(sigma0 + (long)j, 0);
    q.emplace_back(hid, 0);
    q.emplace_back(hid + 1, 0);
    const int64_t n = (int64_t)pk.n;
    std::map<long, std::set<int64_t>> rots;
    for (auto &e : q) rots[e.first].insert(((int64_t)e.second % n + n) % n);
    std::set<std::set<int64_t>> sets;
    for (auto &kv : rots) sets.insert(kv.second);
    const size_t points = A + 2 * L + S + L + TR + 1 + pk.pieces + 2;
    const size_t scalars = pk.committed_queries.size() + pk.advice_queries.size() + pk.fixed_queries.size() + 1 +
                           pk.perm_cols + 3 * S - 1 + 5 * L + TR + sets.size();
    return 48 * points + 32 * scalars;
}

// ---------------------------------------------------------------------- prove

namespace {

struct Query {
    const Fr *poly;     // device coefficients, pk.n long
    HostFr point;
};

}  // namespace

std::vector<uint8_t> plonk_prove(PlonkKey &pk, const RawFr *advice, const RawFr *instance,
                                 const size_t *instance_lens, uint64_t seed)
{
    hipStream_t st = cur_stream();
    const size_t n = pk.n, N = pk.ext_n, A = pk.advice_cols, S = pk.sets;
    const size_t I = pk.instance_cols, L = pk.lookups, TR = pk.trashes;
    Transcript T;

    // every draw of the prover's RNG, in the reference's order
    ChaChaRng rng = ChaChaRng::seed_from_u64(seed);
    const size_t brows = n - pk.u;
    std::vector<HostFr> adv_blinds(A * brows);   // unusable rows u..n of each advice column
    for (auto &b : adv_blinds) b = rng.random_fr();
    // per lookup, the permuted input's then the permuted table's rows u..n
    // (lookup/prover.rs:827-838)
    std::vector<HostFr> lk_in_blinds(L * brows), lk_tab_blinds(L * brows);
    for (size_t l = 0; l < L; ++l) {
        for (size_t j = 0; j < brows; ++j) lk_in_blinds[l * brows + j] = rng.random_fr();
        for (size_t j = 0; j < brows; ++j) lk_tab_blinds[l * brows + j] = rng.random_fr();
    }
    std::vector<HostFr> perm_blinds(S * pk.bf);
    for (auto &b : perm_blinds) b = rng.random_fr();
    std::vector<HostFr> lk_prod_blinds(L * pk.bf);
    for (auto &b : lk_prod_blinds) b = rng.random_fr();
    // vanishing random polynomial: one ChaCha20Rng seed per chunk of the
    // reference prover's thread pool (vanishing/prover.rs:47-66)
    std::vector<size_t> voff, vlen;
    parallelize_chunks(n, pk.rng_threads, voff, vlen);
    std::vector<uint8_t> vseeds(32 * voff.size());
    for (size_t c = 0; c < voff.size(); ++c) rng.fill_bytes(vseeds.data() + 32 * c, 32);
    std::vector<HostFr> limb_blinds(pk.pieces - 1);
    for (auto &b : limb_blinds) b = rng.random_fr();

    // polys (coefficient form, n each): advice | lookup permuted input and
    // table, interleaved per lookup | permutation sets | lookup products |
    // trash | random | h pieces | instance
    const size_t P_ADV = 0, P_LP = A, P_Z = P_LP + 2 * L, P_LZ = P_Z + S, P_TR = P_LZ + L, P_RAND = P_TR + TR,
                 P_H = P_RAND + 1, P_INST = P_H + pk.pieces, NPOLY = P_INST + I;
    // with graphs, the lane's arena at fixed addresses; else one per proof
    LaneProof *lp = nullptr;
    std::unique_ptr<DeviceVec> own_polys, own_fv;
    if (pk.graphs) {
        auto &slot = pk.lane_proofs[current_lane()];
        if (!slot) {
            slot = std::make_unique<LaneProof>();
            slot->polys = new_vec(NPOLY * n);
            slot->fv = new_vec(2 * n);
            // sized once for every captured transform: graphs keep its address
            const size_t widest = std::max({A, S + L, 2 * L, TR, size_t(1)});
            slot->ntt_scratch.alloc(ntt_scratch_elems(pk.k, widest));
        }
        lp = slot.get();
    } else {
        own_polys = new_vec(NPOLY * n);
        own_fv = new_vec(2 * n);
    }
    DeviceVec &polys = lp ? *lp->polys : *own_polys;
    DeviceVec &fvec = lp ? *lp->fv : *own_fv;
    Fr *pd = polys.values.get();
    // commitments of `count` polynomials from index `first` of the arena
    // (after their inverse NTT when `intt`), through the lane's captured
    // stage when graphs are on
    std::vector<RawG1Projective> com;
    auto commit_stage = [&](StageGraph *stage, DeviceVec &vec, size_t first, size_t count, bool intt) {
        Fr *base = vec.values.get() + first * n;
        if (lp) {
            stage_commit(*stage, pk, *lp, base, count, intt, com, st);
            return;
        }
        if (intt) ntt_batch(base, pk.k, count, true, pk.scratch(), st);
        commit(com, pk, vec, first, count);
    };
    // the random polynomial's chunks, drawn on the device while the stream is still empty
    random_fr_streams_device(pd + P_RAND * n, voff.data(), vlen.data(), vseeds.data(), voff.size(), st);

    T.common_scalar(pk.repr);

    // instance columns (compute_instances, prover.rs:1255-1286): Lagrange
    // values (zero beyond each column's length) for the arguments,
    // coefficients for the numerator and the openings; into the transcript
    // go the commitment of each committed column, then the length and values
    // of each other column
    std::unique_ptr<DeviceVec> inst_lag;
    if (I) {
        std::vector<RawFr> iv(I * n);
        std::memset(iv.data(), 0, I * n * 32);
        size_t off = 0;
        for (size_t c = 0; c < I; ++c) {
            const size_t len = instance_lens ? instance_lens[c] : 0;
            MH_CHECK(len <= pk.u, (int)hipErrorInvalidValue, "instance column longer than the usable rows");
            for (size_t i = 0; i < len; ++i) iv[c * n + i] = instance[off + i];
            off += len;
        }
        inst_lag = new_vec(I * n);
        MH_HIP(hipMemcpyAsync(inst_lag->values.get(), iv.data(), I * n * 32, hipMemcpyHostToDevice, st));
        MH_HIP(hipMemcpyAsync(pd + P_INST * n, inst_lag->values.get(), I * n * 32, hipMemcpyDeviceToDevice, st));
        ntt_batch(pd + P_INST * n, pk.k, I, true, pk.scratch(), st);
        MH_HIP(hipStreamSynchronize(st));   // iv goes out of scope
        if (pk.committed) {
            commit(com, pk, polys, P_INST, pk.committed);
            for (auto &c : com) T.common_point(c);
        }
        off = 0;
        for (size_t c = 0; c < I; ++c) {
            const size_t len = instance_lens ? instance_lens[c] : 0;
            if (c >= pk.committed) {
                T.common_scalar(HostFr::from_u64(len));
                for (size_t i = 0; i < len; ++i) {
                    HostFr v;
                    std::memcpy(v.l, instance[off + i].limbs, 32);
                    T.common_scalar(v);
                }
            }
            off += len;
        }
    }

    // advice: Lagrange values (kept for the arguments), coefficients, commitments
    DeviceVec adv_lag;
    adv_lag.lane = current_lane();
    adv_lag.len = A * n;
    adv_lag.values.alloc(A * n);
    Fr *al = adv_lag.values.get();
    MH_HIP(hipMemcpyAsync(al, advice, A * n * 32, hipMemcpyHostToDevice, st));
    // the blinded rows u..n of every column in one strided copy
    if (brows)
        MH_HIP(hipMemcpy2DAsync(al + pk.u, n * 32, adv_blinds.data(), brows * 32, brows * 32, A,
                                hipMemcpyHostToDevice, st));
    MH_HIP(hipMemcpyAsync(pd, al, A * n * 32, hipMemcpyDeviceToDevice, st));
    commit_stage(lp ? &lp->advice : nullptr, polys, P_ADV, A, true);
    for (auto &c : com) T.write_point(c);

    const HostFr theta = T.squeeze();

    // lookups: compressed inputs and tables (theta), permute_expression_pair
    // with the blinds in rows u..n, then the permuted input and table of
    // each lookup committed in that order (lookup/prover.rs:737-1072)
    std::unique_ptr<DeviceVec> ltrace;   // [compressed inputs | tables | permuted inputs | tables]
    if (L) {
        void *lv = nullptr;
        check(midnight_sppark_bls12_381_trace_context_lookup_permute(
            &lv, pk.numctx, &adv_lag, A, inst_lag.get(), I, nullptr, 0, raw(theta), pk.u,
            reinterpret_cast<const RawFr *>(lk_in_blinds.data()), reinterpret_cast<const RawFr *>(lk_tab_blinds.data()),
            brows));
        ltrace.reset(reinterpret_cast<DeviceVec *>(lv));
        const Fr *tr = ltrace->values.get();
        for (size_t l = 0; l < L; ++l) {
            MH_HIP(hipMemcpyAsync(pd + (P_LP + 2 * l) * n, tr + (2 * L + l) * n, n * 32, hipMemcpyDeviceToDevice, st));
            MH_HIP(hipMemcpyAsync(pd + (P_LP + 2 * l + 1) * n, tr + (3 * L + l) * n, n * 32, hipMemcpyDeviceToDevice,
                                  st));
        }
        commit_stage(lp ? &lp->lookups : nullptr, polys, P_LP, 2 * L, true);
        for (auto &c : com) T.write_point(c);
    }

    const HostFr beta = T.squeeze();
    const HostFr gamma = T.squeeze();

    // permutation products from the resident Lagrange advice / instance, then
    // the lookup products (lookup/prover.rs:1075-1340); committed together,
    // permutation sets first
    void *zv = nullptr;
    check(midnight_sppark_bls12_381_trace_context_permutation(
        &zv, pk.numctx, &adv_lag, A, inst_lag.get(), I, raw(beta), raw(gamma),
        reinterpret_cast<const RawFr *>(pk.delta_powers.data()), pk.perm_cols,
        reinterpret_cast<const RawFr *>(perm_blinds.data()), perm_blinds.size(), pk.bf));
    std::unique_ptr<DeviceVec> z(reinterpret_cast<DeviceVec *>(zv));
    MH_HIP(hipMemcpyAsync(pd + P_Z * n, z->values.get(), S * n * 32, hipMemcpyDeviceToDevice, st));
    if (L) {
        void *lzv = nullptr;
        check(midnight_sppark_bls12_381_fr_lookup_products_device(
            &lzv, ltrace.get(), n, L, raw(beta), raw(gamma), reinterpret_cast<const RawFr *>(lk_prod_blinds.data()),
            pk.bf));
        std::unique_ptr<DeviceVec> lz(reinterpret_cast<DeviceVec *>(lzv));
        MH_HIP(hipMemcpyAsync(pd + P_LZ * n, lz->values.get(), L * n * 32, hipMemcpyDeviceToDevice, st));
        MH_HIP(hipStreamSynchronize(st));
    }
    z.reset();
    ltrace.reset();
    commit_stage(lp ? &lp->products : nullptr, polys, P_Z, S + L, true);
    for (auto &c : com) T.write_point(c);

    const HostFr trash = T.squeeze();

    // trash polynomials: each trashcan's constraints folded in the trash
    // challenge on the Lagrange domain, unblinded (trash/prover.rs)
    if (TR) {
        void *tv = nullptr;
        check(midnight_sppark_bls12_381_trace_context_trash(&tv, pk.numctx, &adv_lag, A, inst_lag.get(), I, nullptr,
                                                            0, raw(trash)));
        std::unique_ptr<DeviceVec> tvec(reinterpret_cast<DeviceVec *>(tv));
        MH_HIP(hipMemcpyAsync(pd + P_TR * n, tvec->values.get(), TR * n * 32, hipMemcpyDeviceToDevice, st));
        MH_HIP(hipStreamSynchronize(st));
        tvec.reset();
        commit_stage(lp ? &lp->trash : nullptr, polys, P_TR, TR, true);
        for (auto &c : com) T.write_point(c);
    }

    commit_stage(lp ? &lp->random : nullptr, polys, P_RAND, 1, false);
    T.write_point(com[0]);

    const HostFr y = T.squeeze();

    // numerator on the extended coset of advice | instance | permutation
    // products | lookup products | permuted inputs | permuted tables | trash
    const size_t C_INST = A, C_Z = C_INST + I, C_LZ = C_Z + S, C_LI = C_LZ + L, C_LT = C_LI + L, C_TR = C_LT + L,
                 NCOL = C_TR + TR;
    auto cols = new_vec(NCOL * N);
    Fr *cv = cols->values.get();
    to_extended(cv, pd + P_ADV * n, n, A, pk.ext_k, pk.scratch(), st, true);
    if (I) to_extended(cv + C_INST * N, pd + P_INST * n, n, I, pk.ext_k, pk.scratch(), st, true);
    to_extended(cv + C_Z * N, pd + P_Z * n, n, S + L, pk.ext_k, pk.scratch(), st, true);
    for (size_t l = 0; l < L; ++l) {
        to_extended(cv + (C_LI + l) * N, pd + (P_LP + 2 * l) * n, n, 1, pk.ext_k, pk.scratch(), st, true);
        to_extended(cv + (C_LT + l) * N, pd + (P_LP + 2 * l + 1) * n, n, 1, pk.ext_k, pk.scratch(), st, true);
    }
    if (TR) to_extended(cv + C_TR * N, pd + P_TR * n, n, TR, pk.ext_k, pk.scratch(), st, true);
    MH_HIP(hipStreamSynchronize(st));
    void *hv = nullptr;
    check(numerator_context_eval(&hv, pk.numctx, cols.get(), NCOL, 0, A, C_INST, I, nullptr, 0, raw(beta),
                                 raw(gamma), raw(theta), raw(trash), raw(y), C_Z, S, C_LZ, C_LI, C_LT, L, C_TR, TR,
                                 true, true));
    cols.reset();
    std::unique_ptr<DeviceVec> h(reinterpret_cast<DeviceVec *>(hv));
    // h = numerator / (X^n - 1) on the coset, then coefficients
    // (extended_to_coeff); the numerator is 32 times its value (scaled form),
    // which the division's factors take out
    check(midnight_sppark_bls12_381_fr_device_vec_divide_ntt_zeta(
        h.get(), pk.ext_k, reinterpret_cast<const RawFr *>(pk.t_inv_scaled.data()), pk.t_inv_scaled.size(),
        raw(HostFr::one()), raw(kZeta.sqr()), raw(kZeta)));
    if (limb_blinds.size() <= kMaxArgBlinds) {
        fr_quotient_pieces(pd + P_H * n, h->values.get(), n, pk.pieces,
                           reinterpret_cast<const Fr *>(limb_blinds.data()), limb_blinds.size(), st);
    } else {
        DevBuf<Fr> lb(limb_blinds.size());
        MH_HIP(hipMemcpyAsync(lb.get(), limb_blinds.data(), limb_blinds.size() * 32, hipMemcpyHostToDevice, st));
        fr_quotient_pieces(pd + P_H * n, h->values.get(), n, pk.pieces, lb.get(), st);
        MH_HIP(hipStreamSynchronize(st));   // lb goes out of scope
    }
    MH_HIP(hipStreamSynchronize(st));
    h.reset();
    commit_stage(lp ? &lp->pieces : nullptr, polys, P_H, pk.pieces, false);
    for (auto &c : com) T.write_point(c);

    const HostFr x = T.squeeze();

    // evaluations (write_evals_to_transcript, vanishing / permutation evaluate)
    auto eval1 = [&](const Fr *poly, const HostFr &pt) {
        HostFr v;
        poly_eval_many(&v, poly, n, 1, &pt, 1, st);
        return v;
    };
    // every evaluation the transcript takes after x, in its order, computed
    // by one launch and read back once: advice and fixed queries, the random
    // polynomial, the sigmas, per product set z(x), z(x w) and (all but the
    // last set) z(x w^-(bf+1)), per lookup its product at x and x w, permuted
    // input at x and x w^-1 and permuted table at x (lookup/prover.rs:
    // 1341-1525), per trashcan its polynomial at x
    const Fr *fixed = pk.fixed_coeff->values.get();
    const HostFr x_next = rotate(pk, x, 1), x_last = rotate(pk, x, -(int32_t)(pk.bf + 1)), x_prev = rotate(pk, x, -1);
    {
        std::vector<HostFr> epts;
        std::vector<const Fr *> epolys;
        std::vector<uint32_t> qpoly, qpt;
        auto add = [&](const Fr *poly, const HostFr &pt) {
            size_t pi = 0;
            while (pi < epts.size() && !(epts[pi] == pt)) ++pi;
            if (pi == epts.size()) epts.push_back(pt);
            qpoly.push_back((uint32_t)epolys.size());
            epolys.push_back(poly);
            qpt.push_back((uint32_t)pi);
        };
        for (auto &q : pk.committed_queries) add(pd + (P_INST + q.first) * n, rotate(pk, x, q.second));
        for (auto &q : pk.advice_queries) add(pd + (P_ADV + q.first) * n, rotate(pk, x, q.second));
        for (auto &q : pk.fixed_queries) add(fixed + (size_t)q.first * n, rotate(pk, x, q.second));
        add(pd + P_RAND * n, x);
        for (size_t j = 0; j < pk.perm_cols; ++j) add(pk.sigma_coeff->values.get() + j * n, x);
        for (size_t s = 0; s < S; ++s) {
            add(pd + (P_Z + s) * n, x);
            add(pd + (P_Z + s) * n, x_next);
            if (s + 1 < S) add(pd + (P_Z + s) * n, x_last);
        }
        for (size_t l = 0; l < L; ++l) {
            add(pd + (P_LZ + l) * n, x);
            add(pd + (P_LZ + l) * n, x_next);
            add(pd + (P_LP + 2 * l) * n, x);
            add(pd + (P_LP + 2 * l) * n, x_prev);
            add(pd + (P_LP + 2 * l + 1) * n, x);
        }
        for (size_t t = 0; t < TR; ++t) add(pd + (P_TR + t) * n, x);
        std::vector<HostFr> ev(qpoly.size());
        poly_eval_queries(ev.data(), epolys, n, qpoly.data(), qpt.data(), qpoly.size(), epts.data(), epts.size(),
                          st);
        for (auto &v : ev) T.write_scalar(v);
    }

    // h(X) for the opening: sum_i x^((n-1) i) h_i(X) (vanishing/prover.rs:321-369)
    auto hpoly = new_vec(n);
    {
        const HostFr sf = hpow(x, n - 1);
        std::vector<const Fr *> ps;
        std::vector<HostFr> sc;
        HostFr p = HostFr::one();
        for (size_t i = 0; i < pk.pieces; ++i, p = p * sf) {
            ps.push_back(pd + (P_H + i) * n);
            sc.push_back(p);
        }
        poly_lincomb(hpoly->values.get(), ps, sc, n, HostFr::zero(), st);
    }

    // queries (compute_queries, prover.rs:1993-2080)
    std::vector<Query> queries;
    for (auto &q : pk.advice_queries) queries.push_back({pd + (P_ADV + q.first) * n, rotate(pk, x, q.second)});
    for (auto &q : pk.committed_queries) queries.push_back({pd + (P_INST + q.first) * n, rotate(pk, x, q.second)});
    for (size_t s = 0; s < S; ++s) {
        queries.push_back({pd + (P_Z + s) * n, x});
        queries.push_back({pd + (P_Z + s) * n, x_next});
    }
    for (size_t s = S - 1; s-- > 0;) queries.push_back({pd + (P_Z + s) * n, x_last});
    for (size_t l = 0; l < L; ++l) {   // lookup/prover.rs:1527-1572
        queries.push_back({pd + (P_LZ + l) * n, x});
        queries.push_back({pd + (P_LP + 2 * l) * n, x});
        queries.push_back({pd + (P_LP + 2 * l + 1) * n, x});
        queries.push_back({pd + (P_LP + 2 * l) * n, x_prev});
        queries.push_back({pd + (P_LZ + l) * n, x_next});
    }
    for (size_t t = 0; t < TR; ++t) queries.push_back({pd + (P_TR + t) * n, x});
    for (auto &q : pk.fixed_queries) queries.push_back({fixed + (size_t)q.first * n, rotate(pk, x, q.second)});
    for (size_t j = 0; j < pk.perm_cols; ++j) queries.push_back({pk.sigma_coeff->values.get() + j * n, x});
    queries.push_back({hpoly->values.get(), x});
    queries.push_back({pd + P_RAND * n, x});

    // multi-open (poly/kzg/mod.rs:1978-2570; utils.rs construct_intermediate_sets)
    const HostFr x1 = T.squeeze();
    const HostFr x2 = T.squeeze();
    std::vector<HostFr> points;
    std::vector<const Fr *> cpoly;                 // commitment order
    std::vector<std::vector<size_t>> cpoints;      // point indices per commitment
    for (auto &q : queries) {
        size_t pi = 0;
        while (pi < points.size() && !(points[pi] == q.point)) ++pi;
        if (pi == points.size()) points.push_back(q.point);
        size_t ci = 0;
        while (ci < cpoly.size() && cpoly[ci] != q.poly) ++ci;
        if (ci == cpoly.size()) {
            cpoly.push_back(q.poly);
            cpoints.emplace_back();
        }
        MH_CHECK(std::find(cpoints[ci].begin(), cpoints[ci].end(), pi) == cpoints[ci].end(),
                 (int)hipErrorInvalidValue, "duplicated query");
        cpoints[ci].push_back(pi);
    }
    std::map<std::set<size_t>, size_t> set_index;
    std::vector<size_t> cset(cpoly.size());
    for (size_t c = 0; c < cpoly.size(); ++c) {
        std::set<size_t> ps(cpoints[c].begin(), cpoints[c].end());
        auto it = set_index.find(ps);
        if (it == set_index.end()) it = set_index.emplace(ps, set_index.size()).first;
        cset[c] = it->second;
    }
    const size_t nsets = set_index.size();
    std::vector<std::vector<size_t>> set_points(nsets);
    for (auto &kv : set_index) set_points[kv.second].assign(kv.first.begin(), kv.first.end());
    // q_i = sum_j x1^j p_j over the set's commitments, in commitment order
    auto qv = new_vec(nsets * n);
    for (size_t s = 0; s < nsets; ++s) {
        std::vector<const Fr *> ps;
        std::vector<HostFr> sc;
        HostFr p = HostFr::one();
        for (size_t c = 0; c < cpoly.size(); ++c)
            if (cset[c] == s) {
                ps.push_back(cpoly[c]);
                sc.push_back(p);
                p = p * x1;
            }
        poly_lincomb(qv->values.get() + s * n, ps, sc, n, HostFr::zero(), st);
    }
    // sets in ascending (size, index) order
    std::vector<size_t> order(nsets);
    for (size_t i = 0; i < nsets; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return set_points[a].size() < set_points[b].size(); });
    // f = sum_i x2^i (q_i - r_i) / prod (X - p), i in sorted order
    Fr *f = fvec.values.get(), *work = f + n;   // [f | work]
    {
        std::vector<const Fr *> ps;
        std::vector<HostFr> sc;
        HostFr p = HostFr::one();
        DevBuf<Fr> divided(nsets * n);
        for (size_t oi = 0; oi < nsets; ++oi) {
            const size_t s = order[oi];
            Fr *dst = divided.get() + oi * n;
            MH_HIP(hipMemcpyAsync(dst, qv->values.get() + s * n, n * 32, hipMemcpyDeviceToDevice, st));
            for (size_t pi : set_points[s]) poly_kate_divide(dst, dst, n, points[pi], st);
            ps.push_back(dst);
            sc.push_back(p);
            p = p * x2;
        }
        poly_lincomb(f, ps, sc, n, HostFr::zero(), st);
    }
    commit_stage(lp ? &lp->opening : nullptr, fvec, 0, 1, false);
    T.write_point(com[0]);
    const HostFr x3 = T.squeeze();
    std::vector<HostFr> qe(nsets);
    poly_eval_many(qe.data(), qv->values.get(), n, nsets, &x3, 1, st);   // in set order
    for (size_t oi = 0; oi < nsets; ++oi) T.write_scalar(qe[order[oi]]);
    const HostFr x4 = T.squeeze();
    {
        std::vector<const Fr *> ps;
        std::vector<HostFr> sc;
        HostFr p = HostFr::one();
        for (size_t oi = 0; oi < nsets; ++oi, p = p * x4) {
            ps.push_back(qv->values.get() + order[oi] * n);
            sc.push_back(p);
        }
        ps.push_back(f);
        sc.push_back(p);
        poly_lincomb(work, ps, sc, n, HostFr::zero(), st);
    }
    const HostFr v = eval1(work, x3);
    // pi = (final - v) / (X - x3)
    {
        std::vector<const Fr *> ps{work};
        std::vector<HostFr> sc{HostFr::one()};
        poly_lincomb(work, ps, sc, n, v.neg(), st);
    }
    poly_kate_divide(work, work, n, x3, st);
    MH_HIP(hipMemcpyAsync(f, work, n * 32, hipMemcpyDeviceToDevice, st));
    commit_stage(lp ? &lp->opening : nullptr, fvec, 0, 1, false);
    T.write_point(com[0]);
    MH_HIP(hipStreamSynchronize(st));
    return T.proof;
}

// A proof with no instance values (every instance column of the key empty).
std::vector<uint8_t> plonk_prove(PlonkKey &pk, const RawFr *advice, uint64_t seed)
{
    return plonk_prove(pk, advice, nullptr, nullptr, seed);
}

}  // namespace mh

using namespace mh;

extern "C" RustError midnight_hip_plonk_keygen_with_arguments(
    void **out, uint32_t k, void *srs_context, size_t srs_points, uint32_t degree, const RawFr *fixed,
    size_t fixed_cols, size_t advice_cols, const uint32_t *advice_queries, size_t advice_query_count,
    const uint32_t *fixed_queries, size_t fixed_query_count, const RawColumnRef *perm_columns, size_t perm_cols,
    const uint32_t *perm_mapping, RawGraph gate, const char *description, const RawPlonkArguments *arguments)
{
    if (!out || !srs_context || k < 3 || k > 26 || (fixed_cols && !fixed) || !perm_columns || !perm_mapping ||
        (advice_query_count && !advice_queries) || (fixed_query_count && !fixed_queries))
        return make_error((int)hipErrorInvalidValue, "invalid keygen arguments");
    RawPlonkArguments args{};
    if (arguments) args = *arguments;
    if ((args.lookup_count &&
         (!args.lookup_input_graphs || !args.lookup_table_graphs || !args.lookup_graphs)) ||
        (args.trash_count && (!args.trash_graphs || !args.trash_selectors)))
        return make_error((int)hipErrorInvalidValue, "invalid keygen arguments");
    for (size_t t = 0; t < args.trash_count; ++t)
        if (args.trash_selectors[t] >= fixed_cols)
            return make_error((int)hipErrorInvalidValue, "trash selector column out of range");
    if (args.committed_instances > args.instance_cols || (args.instance_query_count && !args.instance_queries))
        return make_error((int)hipErrorInvalidValue, "invalid instance arguments");
    for (size_t i = 0; i < args.instance_query_count; ++i)
        if (args.instance_queries[2 * i] >= args.instance_cols)
            return make_error((int)hipErrorInvalidValue, "instance query column out of range");
    *out = nullptr;
    MH_TRY
    std::vector<std::pair<uint32_t, int32_t>> aq, fq;
    for (size_t i = 0; i < advice_query_count; ++i) {
        MH_CHECK(advice_queries[2 * i] < advice_cols, (int)hipErrorInvalidValue, "advice query column out of range");
        aq.emplace_back(advice_queries[2 * i], (int32_t)advice_queries[2 * i + 1]);
    }
    for (size_t i = 0; i < fixed_query_count; ++i) {
        MH_CHECK(fixed_queries[2 * i] < fixed_cols, (int)hipErrorInvalidValue, "fixed query column out of range");
        fq.emplace_back(fixed_queries[2 * i], (int32_t)fixed_queries[2 * i + 1]);
    }
    auto pk = plonk_keygen(k, srs_context, srs_points, degree, fixed, fixed_cols, advice_cols, aq, fq, perm_columns,
                           perm_cols, perm_mapping, gate, description ? description : "", args);
    *out = pk.release();
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_hip_plonk_keygen(void **out, uint32_t k, void *srs_context, size_t srs_points,
                                               uint32_t degree, const RawFr *fixed, size_t fixed_cols,
                                               size_t advice_cols, const uint32_t *advice_queries,
                                               size_t advice_query_count, const uint32_t *fixed_queries,
                                               size_t fixed_query_count, const RawColumnRef *perm_columns,
                                               size_t perm_cols, const uint32_t *perm_mapping, RawGraph gate,
                                               const char *description)
{
    return midnight_hip_plonk_keygen_with_arguments(out, k, srs_context, srs_points, degree, fixed, fixed_cols,
                                                    advice_cols, advice_queries, advice_query_count, fixed_queries,
                                                    fixed_query_count, perm_columns, perm_cols, perm_mapping, gate,
                                                    description, nullptr);
}

extern "C" void midnight_hip_plonk_key_free(void *pk)
{
    auto *p = reinterpret_cast<PlonkKey *>(pk);
    if (!p) return;
    try {
        (void)hipStreamSynchronize(lane_stream(p->lane));
    } catch (...) {
    }
    delete p;
}

extern "C" RustError midnight_hip_plonk_vk(const void *pk, RawG1Projective *fixed_commitments,
                                           RawG1Projective *permutation_commitments, RawFr *transcript_repr)
{
    if (!pk) return make_error((int)hipErrorInvalidValue, "null key");
    const auto *p = reinterpret_cast<const PlonkKey *>(pk);
    if (fixed_commitments) std::copy(p->fixed_com.begin(), p->fixed_com.end(), fixed_commitments);
    if (permutation_commitments) std::copy(p->sigma_com.begin(), p->sigma_com.end(), permutation_commitments);
    if (transcript_repr) *transcript_repr = raw(p->repr);
    return ok();
}

extern "C" RustError midnight_hip_plonk_prove_with_instances(void *pk, const RawFr *advice, size_t advice_len,
                                                             const RawFr *instance, const size_t *instance_lens,
                                                             size_t instance_cols, uint64_t rng_seed, uint8_t *proof,
                                                             size_t proof_capacity, size_t *proof_len)
{
    if (!pk || !advice || !proof_len) return make_error((int)hipErrorInvalidValue, "invalid prove arguments");
    auto *p = reinterpret_cast<PlonkKey *>(pk);
    if (advice_len != p->advice_cols * p->n)
        return make_error((int)hipErrorInvalidValue, "advice length does not match the key");
    if (instance_cols != p->instance_cols || (instance_cols && !instance_lens))
        return make_error((int)hipErrorInvalidValue, "instance columns do not match the key");
    size_t total = 0;
    for (size_t c = 0; c < instance_cols; ++c) total += instance_lens[c];
    if (total && !instance) return make_error((int)hipErrorInvalidValue, "null instance values");
    MH_TRY
    // the length is fixed by the key: a short buffer is rejected before any work
    const size_t want = plonk_proof_size(*p);
    *proof_len = want;
    if (!proof || proof_capacity < want) return make_error((int)hipErrorInvalidValue, "proof buffer too small");
    const std::vector<uint8_t> bytes = plonk_prove(*p, advice, instance, instance_lens, rng_seed);
    MH_CHECK(bytes.size() == want, -1, "proof length differs from the key's proof size");
    std::memcpy(proof, bytes.data(), bytes.size());
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_hip_plonk_prove(void *pk, const RawFr *advice, size_t advice_len, uint64_t rng_seed,
                                              uint8_t *proof, size_t proof_capacity, size_t *proof_len)
{
    if (!pk) return make_error((int)hipErrorInvalidValue, "invalid prove arguments");
    const size_t I = reinterpret_cast<const PlonkKey *>(pk)->instance_cols;
    std::vector<size_t> empty(I, 0);
    return midnight_hip_plonk_prove_with_instances(pk, advice, advice_len, nullptr, I ? empty.data() : nullptr, I,
                                                   rng_seed, proof, proof_capacity, proof_len);
}

extern "C" RustError midnight_hip_plonk_set_rng_threads(void *pk, uint32_t threads)
{
    if (!pk || threads == 0) return make_error((int)hipErrorInvalidValue, "invalid prover thread count");
    reinterpret_cast<PlonkKey *>(pk)->rng_threads = threads;
    return ok();
}

extern "C" RustError midnight_hip_plonk_set_graphs(void *pk, bool enable)
{
    if (!pk) return make_error((int)hipErrorInvalidValue, "null key");
    auto *p = reinterpret_cast<PlonkKey *>(pk);
    MH_TRY
    for (auto &lp : p->lane_proofs)
        if (lp) {
            (void)hipStreamSynchronize(cur_stream());
            lp.reset();
        }
    p->graphs = enable;
    return ok();
    MH_CATCH
}

extern "C" RustError midnight_hip_plonk_proof_size(const void *pk, size_t *proof_len)
{
    if (!pk || !proof_len) return make_error((int)hipErrorInvalidValue, "invalid proof size arguments");
    MH_TRY
    *proof_len = plonk_proof_size(*reinterpret_cast<const PlonkKey *>(pk));
    return ok();
    MH_CATCH
}
