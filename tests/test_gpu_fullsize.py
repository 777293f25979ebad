"""GPU parity at the benchmark configurations' full sizes.

- k = 20 proof of the benchmark circuit family (5 advice, 3 fixed, gate
  q(ab + c - d), 8-column copy permutation, 2^22 extended coset): accepted
  by the designated verifier, a tampered proof rejected
  (plonk/prover.rs:1019-1076; verifier restated in tests/plonk_ref.py).
- Proof server on a 2^20-point SRS with k in {12, 14, 16, 18}: every proof
  made in the mix equals the same proof made alone; k <= 14 equal the CPU
  reference prover's bytes, k = 16, 18 accepted by the verifier.
- Batch plan of 256 instances, k = floor(10 + 9u) in [10, 18], over a
  2^18-point context: every instance's coefficients equal the oracle's
  inverse NTT and its MSM the closed form [sum c_i t_i]G.
- 2^20 -> 2^22 coset transform and the vanishing division / extended_to_coeff
  at 2^22 (poly/domain.rs:325-443) against the oracle.
- A circuit with two lookups, a trashcan and a committed instance column at
  2^15 rows: accepted by the designated verifier, rejected tampered or with
  other instance values.
"""
import numpy as np
import pytest

import plonk_ref as PR
import synth_circuit as S

pytestmark = pytest.mark.gpu


def _srs(H, O, n, tau_int):
    powers = H.fr_shifted_prefix_product(np.tile(PR.to_fr(tau_int), (n, 1)), O.fr_from_int(1))
    dv = H.DeviceVec.upload(powers)
    srs = H.MsmContext.from_multiples(dv)
    dv.free()
    return srs


def _fast_key(H, srs, fc):
    return H.PlonkKey(fc.k, srs, S.DEGREE, fc.fixed, S.ADVICE, S.ADVICE_QUERIES, S.FIXED_QUERIES,
                      S.PERM_COLUMNS, fc.mapping, S.gate_graph(), S.DESCRIPTION)


def _circuit_key(H, srs, c):
    fixed = S.mont([v for col in c.cols[5:] for v in col])
    return H.PlonkKey(c.k, srs, S.DEGREE, fixed, S.ADVICE, S.ADVICE_QUERIES, S.FIXED_QUERIES, S.PERM_COLUMNS,
                      S.mapping(c), c.custom_graph(), S.DESCRIPTION)


def _vk(pk):
    fc, pc, rp = pk.vk()
    return fc, pc, PR.to_int(rp)


def _tampered(proof):
    # the first advice evaluation (after the 13 commitments): its low bit
    bad = bytearray(proof)
    bad[13 * 48] ^= 1
    return bytes(bad)


def test_proof_k20_accepted_by_verifier(H, O):
    k = 20
    tau = 0x2f1e3d4c5b6a79880011223344556677
    print("k=20: circuit", flush=True)
    fc = S.FastCircuit(k, seed=2020)
    srs = _srs(H, O, 1 << k, tau)
    print("k=20: keygen", flush=True)
    pk = _fast_key(H, srs, fc)
    assert pk.proof_size() == 15 * 48 + 28 * 32     # 1616 bytes
    print("k=20: prove", flush=True)
    proof = pk.prove(fc.advice, 77)
    assert len(proof) == pk.proof_size()
    vkey = PR.VerifierKey(k, tau)
    assert PR.verify(_vk(pk), proof, vkey)
    assert not PR.verify(_vk(pk), _tampered(proof), vkey)
    assert pk.prove(fc.advice, 77) == proof           # same seed, same bytes
    assert pk.prove(fc.advice, 78) != proof
    with pytest.raises(H.MidnightHipError):
        pk.prove(fc.advice, 77, capacity=len(proof) - 1)   # rejected before proving
    pk.free()
    srs.free()


def test_server_benchmark_mix_on_2p20_srs(H, O):
    tau = 0x51a7e0dd
    print("server: SRS 2^20", flush=True)
    srs = _srs(H, O, 1 << 20, tau)
    small = {k: S.Circuit(k, seed=300 + k) for k in (12, 14)}
    large = {k: S.FastCircuit(k, seed=300 + k) for k in (16, 18)}
    keys, adv = {}, {}
    for k, c in small.items():
        keys[k] = _circuit_key(H, srs, c)
        adv[k] = S.mont([v for col in c.cols[:5] for v in col])
    for k, c in large.items():
        keys[k] = _fast_key(H, srs, c)
        adv[k] = c.advice
    print("server: proofs alone", flush=True)
    alone = {(k, seed): keys[k].prove(adv[k], seed) for k in keys for seed in (1, 2)}
    for k, c in small.items():
        print("server: CPU prover k=%d" % k, flush=True)
        assert alone[(k, 1)] == PR.cpu_prove(PR.CpuKey(c, tau), c.cols[:5], 1), k
    for k in large:
        assert PR.verify(_vk(keys[k]), alone[(k, 1)], PR.VerifierKey(k, tau)), k
    print("server: in the mix", flush=True)
    server = H.ProofServer(12, 64 << 30)
    jobs = [((k, seed), server.submit(keys[k], adv[k], seed))
            for rep in range(2) for k in (12, 14, 12, 16, 12, 14, 18, 12) for seed in (1, 2)]
    # a too small buffer keeps the finished job; the same job is fetched again
    first_key, first_job = jobs[0]
    with pytest.raises(H.MidnightHipError):
        server.wait(first_job, capacity=10)
    for key, jid in jobs:
        assert server.wait(jid)[0] == alone[key], key
    server.free()
    for pk in keys.values():
        pk.free()
    srs.free()


def test_batch_plan_256_instances_k10_to_18(H, O):
    count = 256
    ks = sorted(min(18, max(10, int(10 + 9 * np.random.default_rng(7000 + j).random()))) for j in range(count))
    offs, total = [], 0
    for k in ks:
        offs.append(total)
        total += 1 << k
    ts = O.random_fr(1 << 18, 930)
    dv_t = H.DeviceVec.upload(ts)
    ctx = H.MsmContext.from_multiples(dv_t)
    dv_t.free()
    data = O.random_fr(total, 931)
    dv = H.DeviceVec.upload(data)
    print("batch: plan over %d values" % total, flush=True)
    plan = H.BatchPlan(ctx, dv, offs, ks, H.BatchPlan.GRAPH | H.BatchPlan.AUTOTUNE)
    got = plan.run()
    coeffs = dv.download()
    for i, (o, k) in enumerate(zip(offs, ks)):
        n = 1 << k
        want = O.ifft(data[o:o + n], k)
        assert np.array_equal(coeffs[o:o + n], want), "instance %d (k=%d): coefficients" % (i, k)
        assert O.g1_eq(got[i], O.g1_mul(O.generator(), O.inner_product(want, ts[:n]))), \
            "instance %d (k=%d): MSM" % (i, k)
    plan.free()
    dv.free()
    ctx.free()


def test_coset_and_vanishing_division_2p22(H, O):
    d, t = O.domain(5, 20)
    ek = d.extended_k
    assert ek == 22
    n, N = 1 << d.k, 1 << ek
    coeffs = O.random_fr(n, 2222)
    print("coset 2^22", flush=True)
    src = H.DeviceVec.upload(coeffs)
    ext = H.device_coeff_to_extended_from_device_batch(src, n, 1, ek, d.fr("g_coset"), d.fr("g_coset_inv"))
    want = O.coeff_to_extended(coeffs, d)
    assert np.array_equal(ext.download(), want)
    scale = O.fr_op("mul", d.fr("extended_ifft_divisor"), O.fr_from_int(N))   # = 1, as the prover passes
    v = O.random_fr(N, 2223)
    print("divide 2^22", flush=True)
    got = H.fr_divide_ntt_zeta(v, ek, t, scale, d.fr("g_coset_inv"), d.fr("g_coset"))
    assert np.array_equal(got, O.extended_to_coeff(O.mul_repeating(v, t), d))
    ext.ntt_zeta(ek, True, scale, d.fr("g_coset_inv"), d.fr("g_coset"))     # extended_to_coeff in place
    back = ext.download()
    assert np.array_equal(back[:n], coeffs)
    assert not back[n:].any()
    ext.free()
    src.free()


def test_argument_circuit_k15_accepted_by_verifier(H, O):
    """the circuit with two lookups, a trashcan and a committed instance
    column at 2^15 rows: the GPU proof is accepted by the designated verifier
    (its numerator from the expression trees), rejected tampered or with
    other instance values"""
    k = 15
    tau = 0x1badc0ffee15
    print("args k=15: circuit", flush=True)
    c = S.ArgCircuit(k, seed=1515, committed=1)
    srs = _srs(H, O, 1 << k, tau)
    print("args k=15: keygen", flush=True)
    pk = H.PlonkKey(k, srs, S.DEGREE, S.mont([v for col in c.fixed_values for v in col]), S.ADVICE_X,
                    S.ADVICE_QUERIES_X, S.FIXED_QUERIES_X, S.PERM_COLUMNS_X, S.mapping_x(c), c.custom_graph(),
                    S.DESCRIPTION_X, instance_cols=S.INSTANCE_X, lookups=c.lookup_graphs(), trash=c.trash_graphs(),
                    committed_instances=1, instance_queries=S.INSTANCE_QUERIES_X)
    print("args k=15: prove", flush=True)
    proof = pk.prove(S.mont([v for col in c.advice_values for v in col]), 15, instance=[S.mont(c.instance[0])])
    assert len(proof) == pk.proof_size()
    vkey = PR.VerifierKey(k, tau, args=True, committed=1)
    assert PR.verify(_vk(pk), proof, vkey, c.instance)
    assert not PR.verify(_vk(pk), _tampered(proof), vkey, c.instance)
    assert not PR.verify(_vk(pk), proof, vkey, [[(c.instance[0][0] + 1) % S.R] + c.instance[0][1:]])
    pk.free()
    srs.free()
