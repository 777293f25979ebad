"""GPU proof pipeline (midnight_hip_plonk_keygen / _prove) vs the
reference-algorithm CPU prover: identical proof bytes for the same circuit,
SRS and blinding seed, accepted by the designated verifier.  Stage order and
transcript: plonk/prover.rs:68-575, :769-1010 (see tests/plonk_ref.py)."""
import numpy as np
import pytest

import numerator_graph as G
import plonk_ref as PR
import synth_circuit as S

pytestmark = pytest.mark.gpu


def gpu_key(H, O, c, tau_int, extra_srs=0):
    n = c.n
    powers = H.fr_shifted_prefix_product(np.tile(PR.to_fr(tau_int), (n + extra_srs, 1)), O.fr_from_int(1))
    dv = H.DeviceVec.upload(powers)
    srs = H.MsmContext.from_multiples(dv)
    dv.free()
    fixed = S.mont([v for col in c.cols[5:] for v in col])
    pk = H.PlonkKey(c.k, srs, S.DEGREE, fixed, S.ADVICE, S.ADVICE_QUERIES, S.FIXED_QUERIES, S.PERM_COLUMNS,
                    S.mapping(c), c.custom_graph(), S.DESCRIPTION)
    return pk, srs


@pytest.mark.parametrize("k,seed", [(7, 42), (10, 7), (12, 11), (14, 3)])
def test_gpu_proof_matches_cpu(H, O, k, seed):
    c = S.Circuit(k, seed=3 + k)
    tau = 987654321 + k
    key = PR.CpuKey(c, tau)
    pk, srs = gpu_key(H, O, c, tau, extra_srs=5)
    fc, pc, rp = pk.vk()
    for a, b in zip(fc, key.fixed_com):
        assert O.g1_eq(a, b)
    for a, b in zip(pc, key.sigma_com):
        assert O.g1_eq(a, b)
    assert PR.to_int(rp) == key.repr
    print("proof k=%d" % k, flush=True)
    gpu = pk.prove(S.mont([v for col in c.cols[:5] for v in col]), seed)
    cpu = PR.cpu_prove(key, c.cols[:5], seed)
    assert len(gpu) == len(cpu)
    assert gpu == cpu
    assert PR.verify((fc, pc, PR.to_int(rp)), gpu, key)
    pk.free()
    srs.free()


def test_gpu_proof_rejects_bad_inputs(H, O):
    c = S.Circuit(7, seed=9)
    pk, srs = gpu_key(H, O, c, 5555)
    with pytest.raises(H.MidnightHipError):
        pk.prove(np.zeros((10, 4), dtype=np.uint64), 1)
    with pytest.raises(H.MidnightHipError):
        pk.prove(S.mont([v for col in c.cols[:5] for v in col]), 1, capacity=100)
    pk.free()
    srs.free()


def test_server_results_same_alone_and_in_mix(H, O):
    """Continuous batching: proofs of different k run concurrently on
    several lanes; every proof equals the one made alone (and the CPU
    reference for the small circuits), whatever runs beside it."""
    circuits = [S.Circuit(k, seed=40 + k) for k in (7, 9, 11, 12)]
    keys = [gpu_key(H, O, c, 111 + c.k) for c in circuits]
    adv = [S.mont([v for col in c.cols[:5] for v in col]) for c in circuits]
    alone = {}
    for i, (pk, srs) in enumerate(keys):
        for seed in (1, 2):
            alone[(i, seed)] = pk.prove(adv[i], seed)
    cpu = PR.cpu_prove(PR.CpuKey(circuits[0], 111 + circuits[0].k), circuits[0].cols[:5], 1)
    assert alone[(0, 1)] == cpu
    server = H.ProofServer(4, 8 << 30)
    jobs = []
    for rep in range(3):
        for i, (pk, srs) in enumerate(keys):
            for seed in (1, 2):
                jobs.append(((i, seed), server.submit(pk, adv[i], seed)))
    for key, jid in jobs:
        proof, qms, pms = server.wait(jid)
        assert proof == alone[key], key
        assert qms >= 0 and pms > 0
    server.free()
    # 16 slots: workers on lanes 9 .. 16 too, beyond the lanes callers select
    server = H.ProofServer(16, 8 << 30)
    jobs = [((i, seed), server.submit(pk, adv[i], seed))
            for rep in range(3) for i, (pk, srs) in enumerate(keys) for seed in (1, 2)]
    for key, jid in jobs:
        assert server.wait(jid)[0] == alone[key], key
    server.free()
    with pytest.raises(H.MidnightHipError):
        H.ProofServer(17, 8 << 30)
    # a budget smaller than one proof still serves (one at a time)
    server = H.ProofServer(3, 1)
    jids = [server.submit(keys[1][0], adv[1], 2) for _ in range(4)]
    for jid in jids:
        assert server.wait(jid)[0] == alone[(1, 2)]
    server.free()
    for pk, srs in keys:
        pk.free()
        srs.free()


@pytest.mark.parametrize("rng_threads", [1, 5, 16])
def test_gpu_proof_thread_seeded_random_poly(H, O, rng_threads):
    """The vanishing argument's random polynomial drawn per chunk of an
    rng_threads pool (vanishing/prover.rs:41-75): GPU and CPU prover give the
    same bytes, accepted by the verifier."""
    c = S.Circuit(8, seed=21)
    tau = 24681357
    key = PR.CpuKey(c, tau)
    pk, srs = gpu_key(H, O, c, tau)
    pk.set_rng_threads(rng_threads)
    gpu = pk.prove(S.mont([v for col in c.cols[:5] for v in col]), 5)
    assert gpu == PR.cpu_prove(key, c.cols[:5], 5, rng_threads=rng_threads)
    assert PR.verify((key.fixed_com, key.sigma_com, key.repr), gpu, key)
    with pytest.raises(H.MidnightHipError):
        pk.set_rng_threads(0)
    pk.free()
    srs.free()


@pytest.mark.parametrize("log_n,threads", [(3, 1), (3, 8), (3, 11), (10, 5), (12, 16), (16, 7), (20, 1), (20, 48)])
def test_sample_random_poly_matches_oracle(H, O, log_n, threads):
    """Device ChaCha20Rng draws: chunk c of the parallelize split equals the
    oracle's Fq::random stream of seed c, across candidate blocks, refills
    and chunk boundaries."""
    n = 1 << log_n
    chunks = PR.parallelize_chunks(n, threads)
    rng = O.ChaCha(seed=1000 + log_n + threads)
    seeds = b"".join(rng.fill_bytes(32) for _ in chunks)
    got = H.fr_sample_random_poly(log_n, threads, seeds)
    vals = got.download()
    got.free()
    for ci, (o, ln) in enumerate(chunks):
        if n > 4096 and ci not in (0, 1, len(chunks) - 1):
            continue
        want = O.ChaCha(seed_bytes=seeds[32 * ci:32 * ci + 32]).random_fr(ln)
        assert np.array_equal(vals[o:o + ln], want), (ci, o, ln)
    with pytest.raises(H.MidnightHipError):
        H.fr_sample_random_poly(log_n, threads, seeds[:-32])      # one seed short


def test_gpu_proof_graph_replays_match_eager(H, O):
    """Commitment stages captured in HIP graphs on a lane's first proof and
    replayed after (with another key's proofs on the same lane in between)
    give the same bytes as eager launches and as the CPU prover."""
    c = S.Circuit(9, seed=31)
    tau = 1357924680
    key = PR.CpuKey(c, tau)
    pk, srs = gpu_key(H, O, c, tau)
    other = S.Circuit(8, seed=32)
    pk2, srs2 = gpu_key(H, O, other, tau + 1)
    adv = S.mont([v for col in c.cols[:5] for v in col])
    adv2 = S.mont([v for col in other.cols[:5] for v in col])
    graphs = []
    for seed in (1, 2, 3, 1):
        graphs.append(pk.prove(adv, seed))
        pk2.prove(adv2, seed)
    assert graphs[0] == graphs[3]
    assert graphs[0] == PR.cpu_prove(key, c.cols[:5], 1)
    pk.set_graphs(False)
    eager = [pk.prove(adv, seed) for seed in (1, 2, 3)]
    assert eager == graphs[:3]
    pk.set_graphs(True)
    assert pk.prove(adv, 2) == graphs[1]
    try:
        H.set_lane(3)
        assert pk.prove(adv, 3) == graphs[2]        # a second lane captures its own stages
    finally:
        H.set_lane(0)
    for x in (pk, srs, pk2, srs2):
        x.free()


def gpu_arg_key(H, O, c, tau_int, extra_srs=0):
    """the GPU key of an ArgCircuit (lookups, trashcan, instance column)"""
    n = c.n
    powers = H.fr_shifted_prefix_product(np.tile(PR.to_fr(tau_int), (n + extra_srs, 1)), O.fr_from_int(1))
    dv = H.DeviceVec.upload(powers)
    srs = H.MsmContext.from_multiples(dv)
    dv.free()
    fixed = S.mont([v for col in c.fixed_values for v in col])
    pk = H.PlonkKey(c.k, srs, S.DEGREE, fixed, S.ADVICE_X, S.ADVICE_QUERIES_X, S.FIXED_QUERIES_X,
                    S.PERM_COLUMNS_X, S.mapping_x(c), c.custom_graph(), S.DESCRIPTION_X,
                    instance_cols=S.INSTANCE_X, lookups=c.lookup_graphs(), trash=c.trash_graphs(),
                    committed_instances=c.committed, instance_queries=S.INSTANCE_QUERIES_X)
    return pk, srs


@pytest.mark.parametrize("k,seed,committed", [(7, 5, 0), (10, 8, 0), (12, 2, 0), (7, 4, 1), (11, 9, 1)])
def test_gpu_proof_with_lookups_trash_instance(H, O, k, seed, committed):
    """lookups (two, one of two expressions), a trashcan and an instance
    column tied to advice by a gate and copy constraints, non-committed or
    committed (nb_committed_instances = 1): GPU proof bytes equal to the CPU
    prover's, accepted by the verifier with the instance and rejected with
    other instance values; graph replays equal eager proofs."""
    c = S.ArgCircuit(k, seed=seed, committed=committed)
    tau = 1357911 + k
    key = PR.CpuKey(c, tau)
    pk, srs = gpu_arg_key(H, O, c, tau)
    fc, pc, rp = pk.vk()
    assert PR.to_int(rp) == key.repr
    adv = S.mont([v for col in c.advice_values for v in col])
    inst = [S.mont(col) for col in c.instance]
    print("arg proof k=%d committed=%d" % (k, committed), flush=True)
    gpu = pk.prove(adv, seed, instance=inst)
    assert len(gpu) == pk.proof_size()
    cpu = PR.cpu_prove(key, c.advice_values, seed, instance=c.instance)
    assert gpu == cpu
    vk = (fc, pc, PR.to_int(rp))
    assert PR.verify(vk, gpu, key, c.instance)
    assert not PR.verify(vk, gpu, key, [[(v + 1) % S.R for v in c.instance[0]]])
    assert pk.prove(adv, seed, instance=inst) == gpu          # replayed stages
    pk.set_graphs(False)
    assert pk.prove(adv, seed, instance=inst) == gpu          # eager
    pk.free()
    srs.free()


def test_gpu_proof_lookup_input_missing(H, O):
    """a lookup input absent from its table fails the proof with code -1
    (permute_expression_pair's ConstraintSystemFailure); instance arguments
    that do not fit the key are rejected before any work"""
    c = S.ArgCircuit(7, seed=6)
    pk, srs = gpu_arg_key(H, O, c, 4242)
    adv_cols = [list(x) for x in c.advice_values]
    adv_cols[6][3] = (adv_cols[6][3] + 1) % S.R
    adv = S.mont([v for col in adv_cols for v in col])
    inst = [S.mont(col) for col in c.instance]
    with pytest.raises(H.MidnightHipError) as e:
        pk.prove(adv, 1, instance=inst)
    assert e.value.code == -1
    good = S.mont([v for col in c.advice_values for v in col])
    with pytest.raises(H.MidnightHipError):
        pk.prove(good, 1, instance=inst + inst)                  # two columns for a key of one
    with pytest.raises(H.MidnightHipError):
        pk.prove(good, 1, instance=[S.mont(list(range(c.u + 1)))])   # longer than the usable rows
    # the key still proves after the failures
    key = PR.CpuKey(c, 4242)
    assert pk.prove(good, 1, instance=inst) == PR.cpu_prove(key, c.advice_values, 1, instance=c.instance)
    pk.free()
    srs.free()


def test_server_mixes_argument_circuits(H, O):
    """the proof server with keys of lookups, a trashcan and an instance
    column (committed and not) next to a benchmark-family key: every served
    proof equals the one made alone and the CPU prover's, and verifies"""
    circuits = [S.ArgCircuit(8, seed=21, committed=0), S.ArgCircuit(9, seed=22, committed=1)]
    keys = [gpu_arg_key(H, O, c, 777 + c.k) for c in circuits]
    base = S.Circuit(9, seed=23)
    bkey = gpu_key(H, O, base, 555)
    adv = [S.mont([v for col in c.advice_values for v in col]) for c in circuits]
    inst = [[S.mont(col) for col in c.instance] for c in circuits]
    badv = S.mont([v for col in base.cols[:5] for v in col])
    alone = [pk.prove(adv[i], 5, instance=inst[i]) for i, (pk, _) in enumerate(keys)]
    for i, c in enumerate(circuits):
        cpu_key = PR.CpuKey(c, 777 + c.k)
        assert alone[i] == PR.cpu_prove(cpu_key, c.advice_values, 5, instance=c.instance)
        assert PR.verify((cpu_key.fixed_com, cpu_key.sigma_com, cpu_key.repr), alone[i], cpu_key, c.instance)
    balone = bkey[0].prove(badv, 5)
    server = H.ProofServer(4, 8 << 30)
    jobs = []
    for rep in range(3):
        for i, (pk, _) in enumerate(keys):
            jobs.append((i, server.submit(pk, adv[i], 5, instance=inst[i])))
        jobs.append((-1, server.submit(bkey[0], badv, 5)))
    for i, jid in jobs:
        proof, _, _ = server.wait(jid)
        assert proof == (balone if i < 0 else alone[i]), i
    server.free()
    for pk, srs in keys + [bkey]:
        pk.free()
        srs.free()
