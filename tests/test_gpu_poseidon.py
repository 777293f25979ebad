"""GPU parity of the Poseidon entry points (bls12_381_g1_msm.cu:2820,
:4121, :4192) through the C ABI.

Bar: limb-identical to the oracle (oracle/poseidon.c) -- the chain outputs,
the 256-value witness per hash, and the eight advice columns scattered from
it -- and the reference's errors for rejected arguments.
"""
import random

import numpy as np
import pytest

from test_oracle_poseidon import R, flat_mont, mont, packed_layout, params

pytestmark = pytest.mark.gpu


def _chain_args(seed):
    rnd = random.Random(seed)
    return [mont([rnd.randrange(R)])[0] for _ in range(3)] + [mont([3])[0]]


@pytest.mark.parametrize("length,seed", [(1, 1), (3, 2), (64, 3)])
def test_hash_chain_matches_oracle(H, O, length, seed):
    m, k = flat_mont(*params(seed))
    args = _chain_args(seed)
    got = H.fr_poseidon_hash_chain(length, *args, m, k)
    assert np.array_equal(got, O.poseidon_hash_chain(length, *args, m, k))


@pytest.mark.parametrize("length,seed", [(1, 4), (5, 5)])
def test_hash_chain_witness_matches_oracle(H, O, length, seed):
    m, k = flat_mont(*params(seed))
    args = _chain_args(seed)
    got = H.fr_poseidon_hash_chain_witness(length, *args, m, k)
    want = O.poseidon_hash_chain(length, *args, m, k, witness=True)
    assert got.shape == (length * 256, 4)
    assert np.array_equal(got, want)
    # slot 253 of each hash is the chain output
    outs = H.fr_poseidon_hash_chain(length, *args, m, k)
    assert np.array_equal(got[253::256], outs)


def test_hash_chain_edge_inputs(H, O):
    """zero and p - 1 inputs, an all-zero parameter set, and an identity MDS"""
    mds, rc = params(6)
    m, k = flat_mont(mds, rc)
    zero, top, three = mont([0])[0], mont([R - 1])[0], mont([3])[0]
    for args in ([zero, zero, zero, zero], [top, top, top, three]):
        assert np.array_equal(H.fr_poseidon_hash_chain(4, *args, m, k), O.poseidon_hash_chain(4, *args, m, k))
    ident = mont([1, 0, 0, 0, 1, 0, 0, 0, 1])
    zeros = np.zeros((204, 4), dtype=np.uint64)
    args = [mont([5])[0], mont([6])[0], mont([7])[0], three]
    for mm, kk in ((ident, k), (m, zeros), (ident, zeros)):
        got = H.fr_poseidon_hash_chain_witness(2, *args, mm, kk)
        assert np.array_equal(got, O.poseidon_hash_chain(2, *args, mm, kk, witness=True))


def test_hash_chain_rejects_bad_parameters(H):
    m, k = flat_mont(*params(7))
    args = _chain_args(7)
    # len == 0 succeeds without looking at anything else
    assert H.fr_poseidon_hash_chain(0, *args, m[:8], k).shape == (0, 4)
    for fn, msg in ((H.fr_poseidon_hash_chain, "hash-chain"), (H.fr_poseidon_hash_chain_witness, "witness")):
        with pytest.raises(H.MidnightHipError, match=msg):
            fn(2, *args, m[:8], k)
        with pytest.raises(H.MidnightHipError, match=msg):
            fn(2, *args, m, k[:203])


@pytest.mark.parametrize("hashes,domain", [(1, 64), (5, 256), (24, 1024)])
def test_advice_from_witness_matches_oracle(H, O, hashes, domain):
    mds, rc = params(8)
    m, k = flat_mont(mds, rc)
    args = _chain_args(hashes)
    wit = H.fr_poseidon_hash_chain_witness(hashes, *args, m, k)
    rnd = random.Random(hashes)
    metadata = np.concatenate([mont([rnd.randrange(R) for _ in range(4)]), np.stack(args), k[:3]])
    layout = packed_layout()
    dv = H.fr_poseidon_advice_from_witness(wit, hashes, domain, metadata, layout)
    got = dv.download()
    want = O.poseidon_advice_from_witness(wit, hashes, domain, metadata, layout)
    assert got.shape == (8 * domain, 4)
    assert np.array_equal(got, want)


def test_advice_from_witness_rejects_bad_arguments(H):
    m, k = flat_mont(*params(9))
    wit = H.fr_poseidon_hash_chain_witness(2, *_chain_args(9), m, k)
    metadata = np.zeros((11, 4), dtype=np.uint64)
    layout = packed_layout()
    bad_col = layout.copy()
    bad_col[3] = (8 << 8) | 1
    bad_row = layout.copy()
    bad_row[0] = (3 << 8) | 255
    for w, hc, dom, meta, lay in ((wit, 2, 84, metadata, layout),          # 2 + 41 * 2 >= 84
                                  (wit, 0, 256, metadata, layout),
                                  (wit, 2, 256, metadata[:10], layout),
                                  (wit, 2, 256, metadata, layout[:127]),
                                  (wit, 2, 256, metadata, bad_col),
                                  (wit, 2, 256, metadata, bad_row)):      # 2 + 41 + 21 + 255 >= 256
        with pytest.raises(H.MidnightHipError, match="Poseidon advice"):
            H.fr_poseidon_advice_from_witness(w, hc, dom, meta, lay)
