"""CPU side of the proof pipeline: the reference-algorithm CPU prover's
proofs are accepted by the designated verifier (which knows the SRS
trapdoor and checks the multi-open as a G1 equation), and tampered proofs
or a wrong witness are rejected; the oracle's polynomial helpers are pinned
against direct integer arithmetic; the transcript hash is pinned by
hashlib's Blake2b (RFC 7693)."""
import random

import numpy as np
import pytest

import oracle as O
import plonk_ref as PR
import synth_circuit as S

R = S.R


def test_poly_eval_and_kate_division():
    rnd = random.Random(5)
    n = 37
    coeffs = [rnd.randrange(R) for _ in range(n)]
    x = rnd.randrange(R)
    assert PR.to_int(O.poly_eval(S.mont(coeffs), PR.to_fr(x))) == sum(c * pow(x, i, R) for i, c in enumerate(coeffs)) % R
    b = rnd.randrange(R)
    q = S.ints(O.kate_division(S.mont(coeffs), PR.to_fr(b)))
    assert q[-1] == 0
    ab = sum(c * pow(b, i, R) for i, c in enumerate(coeffs)) % R
    # q(X) (X - b) + a(b) == a(X)
    prod = [0] * n
    for i in range(n - 1):
        prod[i + 1] = (prod[i + 1] + q[i]) % R
        prod[i] = (prod[i] - b * q[i]) % R
    prod[0] = (prod[0] + ab) % R
    assert prod == coeffs


@pytest.fixture(scope="module")
def setup():
    c = S.Circuit(7, seed=3)
    key = PR.CpuKey(c, 123456789)
    return c, key


def test_cpu_proof_verifies(setup):
    c, key = setup
    proof = PR.cpu_prove(key, c.cols[:5], 42)
    assert len(proof) == 48 * (5 + 3 + 1 + 4) + 32 * (5 + 3 + 1 + 8 + 8) + 48 + 32 * 3 + 48
    vk = (key.fixed_com, key.sigma_com, key.repr)
    assert PR.verify(vk, proof, key)
    # the same seed gives the same bytes; another seed other blinding
    assert PR.cpu_prove(key, c.cols[:5], 42) == proof
    assert PR.cpu_prove(key, c.cols[:5], 43) != proof


def test_tampered_proofs_rejected(setup):
    c, key = setup
    proof = PR.cpu_prove(key, c.cols[:5], 7)
    vk = (key.fixed_com, key.sigma_com, key.repr)
    for pos in (10, 48 * 6 + 3, 48 * 13 + 5, len(proof) - 60, len(proof) - 5):
        bad = bytearray(proof)
        bad[pos] ^= 1
        assert not PR.verify(vk, bytes(bad), key), pos
    assert not PR.verify(vk, proof[:-1], key)
    assert not PR.verify(vk, proof + b"\x00", key)


def test_wrong_witness_rejected(setup):
    c, key = setup
    vk = (key.fixed_com, key.sigma_com, key.repr)
    cols = [list(x) for x in c.cols[:5]]
    cols[0][3] = (cols[0][3] + 1) % R           # gate fails at row 3
    assert not PR.verify(vk, PR.cpu_prove(key, cols, 42), key)
    cols = [list(x) for x in c.cols[:5]]
    j, i = next(cell for cyc in c.cycles for cell in cyc if cell[0] == 4)
    cols[j][i] = (cols[j][i] + 1) % R          # a copy constraint fails
    assert not PR.verify(vk, PR.cpu_prove(key, cols, 42), key)


def test_transcript_blake2b():
    import hashlib
    t = PR.Transcript()
    t.absorb(b"abc")
    ch = t.squeeze()
    h = hashlib.blake2b(digest_size=64, key=PR.KEY)
    h.update(b"\x01abc\x00")
    assert ch == int.from_bytes(h.digest(), "little") % R


def test_parallelize_chunks_reference_example():
    """utils/arithmetic.rs:145-151: 40 items on 12 threads are cut into
    4, 4, 4, 4, 3, 3, 3, 3, 3, 3, 3, 3."""
    chunks = PR.parallelize_chunks(40, 12)
    assert [ln for _, ln in chunks] == [4] * 4 + [3] * 8
    assert [o for o, _ in chunks] == [0, 4, 8, 12, 16, 19, 22, 25, 28, 31, 34, 37]


@pytest.mark.parametrize("n,threads", [(128, 1), (128, 5), (128, 16), (1024, 5), (1024, 12), (5, 8), (8, 8),
                                       (1 << 20, 48), (1 << 20, 7)])
def test_sample_random_seeds_cover_the_chunks(n, threads):
    """Every chunk parallelize hands out gets its own seed: the offsets of
    vanishing/prover.rs:50-66, in seed order, are the chunk offsets in row
    order, the chunks partition the rows and differ in length by at most 1."""
    chunks = PR.parallelize_chunks(n, threads)
    assert PR.sample_random_seed_offsets(n, threads) == [o for o, _ in chunks]
    assert len(chunks) == min(n, threads)
    pos = 0
    for o, ln in chunks:
        assert o == pos and ln >= 1
        pos += ln
    assert pos == n
    assert max(ln for _, ln in chunks) - min(ln for _, ln in chunks) <= 1


def test_sample_random_draw_order():
    """The polynomial of a pool of T threads takes T 32-byte seeds from the
    prover's RNG: the draws after it (the quotient-limb blinds) start 8 T
    words later; with one thread the polynomial is one stream."""
    n = 64
    for threads in (1, 3, 16):
        rng = O.ChaCha(seed=5)
        poly = PR.sample_random(rng, n, threads)
        assert poly.shape == (n, 4)
        after = rng.random_fr(2)
        ref = O.ChaCha(seed=5)
        seeds = [ref.fill_bytes(32) for _ in range(threads)]
        assert np.array_equal(ref.random_fr(2), after)
        o, ln = PR.parallelize_chunks(n, threads)[-1]
        assert np.array_equal(poly[o:], O.ChaCha(seed_bytes=seeds[-1]).random_fr(ln))
    one = PR.sample_random(O.ChaCha(seed=9), n, 1)
    assert np.array_equal(one, O.ChaCha(seed_bytes=O.ChaCha(seed=9).fill_bytes(32)).random_fr(n))


def test_fast_circuit_satisfies_its_constraints():
    """The numpy-built circuit of the family: the gate q (a b + c - d) is
    zero on every row and every copy cycle holds one value."""
    fc = S.FastCircuit(9, seed=4)
    n = fc.n
    cols = np.concatenate([fc.advice, fc.fixed]).reshape(8, n, 4)
    a, b, c, d, e, q, s1, s2 = (S.ints(cols[i]) for i in range(8))
    for i in range(n):
        assert q[i] * (a[i] * b[i] + c[i] - d[i]) % R == 0, i
    m = fc.mapping.reshape(8, n, 2)
    moved = 0
    for j in range(8):
        for i in range(n):
            jj, ii = int(m[j, i, 0]), int(m[j, i, 1])
            vals = [a, b, c, d, e, q, s1, s2]
            assert vals[j][i] == vals[jj][ii], (j, i)
            # the mapping is a permutation: the target maps back in a cycle of two
            assert (int(m[jj, ii, 0]), int(m[jj, ii, 1])) == (j, i)
            moved += (jj, ii) != (j, i)
    assert moved == 2 * (fc.u // 4)


def test_cpu_proof_with_thread_seeded_random_poly(setup):
    """A proof whose random polynomial follows a 5-thread pool is accepted
    and differs from the one-thread proof (every later draw moves)."""
    c, key = setup[:2]
    p1 = PR.cpu_prove(key, c.cols[:5], 3)
    p5 = PR.cpu_prove(key, c.cols[:5], 3, rng_threads=5)
    assert p1 != p5
    assert PR.verify((key.fixed_com, key.sigma_com, key.repr), p5, key)


def test_cpu_prover_on_fast_circuit():
    """The CPU reference prover's key built from the numpy circuit (sigma
    from the mapping by vectorised products): its proof verifies, and a
    witness that breaks one copy cycle does not."""
    fc = S.FastCircuit(8, seed=12)
    key = PR.CpuKey(fc, 97531)
    vk = (key.fixed_com, key.sigma_com, key.repr)
    proof = PR.cpu_prove(key, fc.advice, 4)
    assert PR.verify(vk, proof, key)
    bad = fc.advice.copy().reshape(5, fc.n, 4)
    row = int(np.nonzero((fc.mapping.reshape(8, fc.n, 2)[4, :, 0] == 0))[0][0])   # a copied cell of e
    bad[4, row] = O.fr_op("add", bad[4, row], O.fr_from_int(1))
    assert not PR.verify(vk, PR.cpu_prove(key, bad.reshape(-1, 4), 4), key)


# ---- lookups, a trashcan and an instance column (plonk/lookup, plonk/trash,
# compute_instances of plonk/prover.rs:1237-1330)

@pytest.fixture(scope="module")
def arg_setup():
    c = S.ArgCircuit(7, seed=3)
    key = PR.CpuKey(c, 24681357)
    return c, key


def test_arg_circuit_satisfies_its_constraints(arg_setup):
    """the witness the verifier's numerator is built against: gates, copy
    cycles (instance cells included), both lookups and the trashcan hold on
    every usable row"""
    c, _ = arg_setup
    a, b, cc, d, e, f, g = c.advice_values
    q, s1, s2, t1, t2, qp, qt = c.fixed_values
    p = c.instance[0] + [0] * (c.n - len(c.instance[0]))
    cols = c.cols + [p]
    for i in range(c.u):
        assert q[i] * (a[i] * b[i] + cc[i] - d[i]) % R == 0
        assert qp[i] * (a[i] - p[i]) % R == 0
        assert qt[i] * (d[i] - a[i] * b[i] - cc[i]) % R == 0
    for cyc in c.cycles:
        assert len({cols[j][i] for j, i in cyc}) == 1
    table0 = {(t1[i], t2[i]) for i in range(c.u)}
    assert all((f[i], g[i]) in table0 for i in range(c.u))
    table1 = set(s1[:c.u])
    assert all(q[i] * e[i] % R in table1 for i in range(c.u))


def test_arg_circuit_cpu_proof_verifies(arg_setup):
    c, key = arg_setup
    proof = PR.cpu_prove(key, c.advice_values, 42, instance=c.instance)
    A, F, L, TR, Sn = S.ADVICE_X, S.FIXED_X, 2, 1, 3
    # commitments: advice, permuted pairs, permutation sets, lookup products,
    # trash, random, pieces, f, pi; evaluations: advice, fixed, random,
    # sigmas, sets, lookups, trash, then one q(x3) per point set
    points = A + 2 * L + Sn + L + TR + 1 + 4 + 2
    scalars = A + F + 1 + 9 + (3 * Sn - 1) + 5 * L + TR
    assert (len(proof) - 48 * points - 32 * scalars) % 32 == 0
    vk = (key.fixed_com, key.sigma_com, key.repr)
    assert PR.verify(vk, proof, key, c.instance)
    assert PR.cpu_prove(key, c.advice_values, 42, instance=c.instance) == proof
    # other instance values: rejected (the gate qp (a - p) and the copies fail)
    other = [[(c.instance[0][0] + 1) % R] + c.instance[0][1:]]
    assert not PR.verify(vk, proof, key, other)
    assert not PR.verify(vk, proof, key, [c.instance[0][:-1]])
    for pos in (10, 48 * 8 + 3, len(proof) - 60):
        bad = bytearray(proof)
        bad[pos] ^= 1
        assert not PR.verify(vk, bytes(bad), key, c.instance), pos


def test_arg_circuit_wrong_witness(arg_setup):
    c, key = arg_setup
    vk = (key.fixed_com, key.sigma_com, key.repr)
    adv = [list(x) for x in c.advice_values]
    adv[5][4] = (adv[5][4] + 1) % R             # lookup input (f, g) not in (t1, t2)
    with pytest.raises(ValueError):
        PR.cpu_prove(key, adv, 42, instance=c.instance)
    adv = [list(x) for x in c.advice_values]
    adv[3][5] = (adv[3][5] + 1) % R             # gate and trashcan (row 5: qt = 1) fail
    assert not PR.verify(vk, PR.cpu_prove(key, adv, 42, instance=c.instance), key, c.instance)
    # the instance passed to the prover differs from the cells it is tied to
    wrong = [[(v + 1) % R for v in c.instance[0]]]
    assert not PR.verify(vk, PR.cpu_prove(key, c.advice_values, 42, instance=wrong), key, wrong)


def test_arg_circuit_committed_instance():
    """nb_committed_instances = 1: the instance column's commitment goes
    into the transcript instead of its values, its query is evaluated after
    x (before the advice) and opened; the verifier computes the commitment
    from the public values"""
    c = S.ArgCircuit(7, seed=3, committed=1)
    key = PR.CpuKey(c, 24681357)
    proof = PR.cpu_prove(key, c.advice_values, 42, instance=c.instance)
    vk = (key.fixed_com, key.sigma_com, key.repr)
    assert PR.verify(vk, proof, key, c.instance)
    other = [[(c.instance[0][0] + 1) % R] + c.instance[0][1:]]
    assert not PR.verify(vk, proof, key, other)
    # the same witness proved with the column not committed: the advice
    # commitments agree, theta and everything after it differ
    c0 = S.ArgCircuit(7, seed=3, committed=0)
    key0 = PR.CpuKey(c0, 24681357)
    proof0 = PR.cpu_prove(key0, c0.advice_values, 42, instance=c0.instance)
    assert len(proof) == len(proof0) + 32
    assert proof[:48 * 7] == proof0[:48 * 7] and proof[48 * 7:48 * 8] != proof0[48 * 7:48 * 8]
    assert not PR.verify(vk, proof0, key, c.instance)
