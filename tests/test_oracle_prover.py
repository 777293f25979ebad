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
