"""Pins the oracle's Poseidon restatement (oracle/poseidon.c).

- the permutation against a textbook HADES round function written here in
  Python integers (add round key, S-box x^5 on all lanes in the 4 + 4 full
  rounds and on the last lane in the 60 partial ones, MDS), the form of
  poseidon_cpu.rs:53-67,126-130 rather than the reference kernel's folding of
  each round's key into the previous MDS step;
- the permutation against its inverse (inverse MDS, x^(1/5), subtract key);
- the 128-value witness against the textbook trace: per full round the cubes
  of the S-box inputs and the next S-box inputs, per batch of six partial
  rounds five last-lane values and one whole state (poseidon_cpu.rs:164-188);
- the chain against Python chaining of the textbook permutation
  (poseidon_cpu.rs:222-240);
- the advice scatter against the columns built directly while permuting,
  in the order of hash_chain_circuit_advice_columns (poseidon_cpu.rs:246-279)
  with the layout of poseidon_permutation_witness_layout
  (poseidon_chip.rs:108-128).
"""
import random

import numpy as np
import pytest

import numerator_graph as G

R = G.FR_MODULUS
ROUNDS = 68


def full(r):
    return r < 4 or r >= 64


def mont(values):
    return G.to_montgomery_limbs(values)


def ints(arr):
    return G.from_montgomery_limbs(np.asarray(arr).reshape(-1, 4))


def params(seed):
    rnd = random.Random(seed)
    mds = [[rnd.randrange(R) for _ in range(3)] for _ in range(3)]
    rc = [[rnd.randrange(R) for _ in range(3)] for _ in range(ROUNDS)]
    return mds, rc


def flat_mont(mds, rc):
    return mont([x for row in mds for x in row]), mont([x for row in rc for x in row])


def hades_trace(state, mds, rc):
    """S-box inputs of every round (after its round key) and the output"""
    trace = []
    for r in range(ROUNDS):
        state = [(s + k) % R for s, k in zip(state, rc[r])]
        trace.append(state)
        if full(r):
            state = [pow(s, 5, R) for s in state]
        else:
            state = state[:2] + [pow(state[2], 5, R)]
        state = [sum(mds[i][j] * state[j] for j in range(3)) % R for i in range(3)]
    trace.append(state)
    return trace


def hades(state, mds, rc):
    return hades_trace(state, mds, rc)[-1]


def mat_inv3(m):
    det = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
           - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
           + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0])) % R
    di = pow(det, -1, R)
    cof = [[(m[(j + 1) % 3][(i + 1) % 3] * m[(j + 2) % 3][(i + 2) % 3]
             - m[(j + 1) % 3][(i + 2) % 3] * m[(j + 2) % 3][(i + 1) % 3]) % R for j in range(3)] for i in range(3)]
    return [[cof[i][j] * di % R for j in range(3)] for i in range(3)]


def hades_inverse(state, mds, rc):
    mi = mat_inv3(mds)
    root5 = pow(5, -1, R - 1)
    for r in reversed(range(ROUNDS)):
        state = [sum(mi[i][j] * state[j] for j in range(3)) % R for i in range(3)]
        if full(r):
            state = [pow(s, root5, R) for s in state]
        else:
            state = state[:2] + [pow(state[2], root5, R)]
        state = [(s - k) % R for s, k in zip(state, rc[r])]
    return state


def textbook_witness(trace):
    w = []
    for r in range(ROUNDS):
        if full(r):
            w += [pow(x, 3, R) for x in trace[r]] + list(trace[r + 1])
        elif (r - 4) % 6 < 5:
            w.append(trace[r + 1][2])
        else:
            w += list(trace[r + 1])
    return w


def standard_layout():
    layout, offset = [], 1
    for _ in range(4):
        layout += [(col, offset) for col in range(3, 6)]
        offset += 1
        layout += [(col, offset) for col in range(3)]
    for _ in range(10):
        layout += [(col, offset) for col in range(3, 8)]
        offset += 1
        layout += [(col, offset) for col in range(3)]
    for _ in range(4):
        layout += [(col, offset) for col in range(3, 6)]
        offset += 1
        layout += [(col, offset) for col in range(3)]
    assert len(layout) == 128
    return layout


def packed_layout():
    return np.array([(c << 8) | r for c, r in standard_layout()], dtype=np.uint16)


def test_permutation_matches_textbook_and_inverse(O):
    for seed in range(3):
        mds, rc = params(seed)
        m, k = flat_mont(mds, rc)
        rnd = random.Random(100 + seed)
        for state in ([0, 0, 0], [1, 2, 3], [rnd.randrange(R) for _ in range(3)], [R - 1, R - 1, R - 1]):
            got = ints(O.poseidon_permutation(mont(state), m, k))
            assert got == hades(state, mds, rc)
            assert hades_inverse(got, mds, rc) == state


def test_permutation_witness_matches_textbook_trace(O):
    mds, rc = params(7)
    m, k = flat_mont(mds, rc)
    state = [5, 6, 7]
    out, w = O.poseidon_permutation(mont(state), m, k, witness=True)
    trace = hades_trace(state, mds, rc)
    assert ints(w) == textbook_witness(trace)
    # the last three witness values are the permutation's output
    assert ints(w)[125:128] == ints(out) == trace[-1]


def test_hash_chain_and_witness(O):
    mds, rc = params(11)
    m, k = flat_mont(mds, rc)
    rnd = random.Random(5)
    initial, tail0, tail1 = (rnd.randrange(R) for _ in range(3))
    n = 4
    args = [mont([v])[0] for v in (initial, tail0, tail1, 3)]
    out = ints(O.poseidon_hash_chain(n, *args, m, k))
    wit = ints(O.poseidon_hash_chain(n, *args, m, k, witness=True))
    prev, want, want_w = initial, [], []
    for _ in range(n):
        t1 = hades_trace([prev, tail0, 3], mds, rc)
        s = list(t1[-1])
        s[0] = (s[0] + tail1) % R
        t2 = hades_trace(s, mds, rc)
        want_w += textbook_witness(t1) + textbook_witness(t2)
        prev = t2[-1][0]
        want.append(prev)
    assert out == want
    assert wit == want_w
    # the chain output is witness slot 253 of each hash
    assert [wit[i * 256 + 253] for i in range(n)] == out


def test_advice_from_witness_matches_direct_columns(O):
    mds, rc = params(13)
    m, k = flat_mont(mds, rc)
    rnd = random.Random(17)
    initial, tail0, tail1 = (rnd.randrange(R) for _ in range(3))
    prefix = [rnd.randrange(R) for _ in range(3)]
    binding = rnd.randrange(R)
    hashes, domain = 3, 128
    wit = O.poseidon_hash_chain(hashes, *[mont([v])[0] for v in (initial, tail0, tail1, 3)], m, k, witness=True)
    metadata = mont(prefix + [binding, initial, tail0, tail1, 3] + rc[0])
    got = ints(O.poseidon_advice_from_witness(wit, hashes, domain, metadata, packed_layout()))

    # hash_chain_circuit_advice_columns: columns filled while permuting
    cols = [[0] * domain for _ in range(8)]
    layout = standard_layout()

    def permute_into(state, row_start):
        for c in range(3):
            cols[c][row_start] = state[c]
            cols[c][row_start + 1] = (state[c] + rc[0][c]) % R
        trace = hades_trace(state, mds, rc)
        for (c, r), v in zip(layout, textbook_witness(trace)):
            cols[c][row_start + r] = v
        return list(trace[-1])

    prev = initial
    for h in range(hashes):
        row = 2 + 41 * h
        first = permute_into([prev, tail0, 3], row)
        s = [(first[0] + tail1) % R, first[1], first[2]]
        cols[0][row + 20], cols[1][row + 20], cols[2][row + 20] = s[0], first[0], tail1
        prev = permute_into(s, row + 21)[0]
    # prefix rows the caller assigns
    cols[0][0], cols[1][0], cols[2][0], cols[0][1] = prefix[0], prefix[1], prefix[2], binding
    assert got == [v for col in cols for v in col]
    # columns past the layout's last (7) stay empty beyond the used rows
    assert all(v == 0 for v in got[7 * domain + 2 + 41 * hashes:])


@pytest.mark.parametrize("hashes", [1, 2])
def test_permutation_is_deterministic_per_parameter_set(O, hashes):
    """A changed round constant changes every later output (no constant is skipped)."""
    mds, rc = params(21)
    m, k = flat_mont(mds, rc)
    args = [mont([v])[0] for v in (9, 8, 7, 3)]
    base = ints(O.poseidon_hash_chain(hashes, *args, m, k))
    for r in (0, 1, 3, 4, 33, 63, 64, 67):
        for c in range(3):
            rc2 = [list(row) for row in rc]
            rc2[r][c] = (rc2[r][c] + 1) % R
            _, k2 = flat_mont(mds, rc2)
            assert ints(O.poseidon_hash_chain(hashes, *args, m, k2))[-1] != base[-1]
