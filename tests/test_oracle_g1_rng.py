"""Pins the oracle's G1 group law, encodings and RNG (CPU only).

- 2G Montgomery limbs: vendor/midnight-curves-0.3.1/src/bls12_381/g1.rs:1589-1622
- [r]G = O, P + (-P) = O, on-curve results
- affine group law against Python big integers (chord-and-tangent)
- ZCash compressed round trip, and the reference SRS data file
  (transient-crypto/static/bls_midnight_2p14 re-encoded in
  tests/golden/srs_2p14.g1c): g[0] is the generator, every point on-curve
- ChaCha20 block function: RFC 7539 section 2.3.2 test vector; the
  ChaCha20Rng stream: RFC 7539 appendix A.1 keystream vectors #1, #2; the
  PCG32 step and output of seed_from_u64: the pcg32 demo outputs
"""
import os
import random

import numpy as np
import pytest

P_MOD = 0x1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab
R_MOD = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
REF_SRS = "/root/reference/transient-crypto/static/bls_midnight_2p14"


def fp_int(a):
    return sum(int(x) << (64 * i) for i, x in enumerate(a)) * pow(1 << 384, -1, P_MOD) % P_MOD


def affine_int(O, j):
    a = O.to_affine(j)
    if not a.any():
        return None
    return fp_int(a[:6]), fp_int(a[6:])


def py_add(p, q):
    if p is None:
        return q
    if q is None:
        return p
    (x1, y1), (x2, y2) = p, q
    if x1 == x2:
        if (y1 + y2) % P_MOD == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, P_MOD) % P_MOD
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P_MOD) % P_MOD
    x3 = (lam * lam - x1 - x2) % P_MOD
    return x3, (lam * (x1 - x3) - y1) % P_MOD


def test_double_generator_kat(O):
    g = O.generator()
    gj = O.g1j()
    O.lib().orc_g1j_from_affine(O.P(gj), O.P(g))
    d = O.g1j()
    O.lib().orc_g1j_dbl(O.P(d), O.P(gj))
    want = [0x53e978ce58a9ba3c, 0x3ea0583c4f3d65f9, 0x4d20bb47f0012960, 0xa54c664ae5b2b5d9, 0x26b552a39d7eb21f,
            0x8895d26e68785, 0x70110b3298293940, 0xda33c5393f1f6afc, 0xb86edfd16a5aa785, 0xaec6d1c9e7b1c895,
            0x25cfc2b522d11720, 0x6361c83f8d09b15]
    assert list(O.to_affine(d)) == want


def test_group_order_and_inverse(O):
    g = O.generator()
    r_as_fr = O.fr()   # r = 0 mod r: use (r-1)G + G
    rm1 = O.fr_op("neg", O.fr_from_int(1))
    a = O.g1_mul(g, rm1)
    gj = O.g1j()
    O.lib().orc_g1j_from_affine(O.P(gj), O.P(g))
    s = O.g1j()
    O.lib().orc_g1j_add(O.P(s), O.P(a), O.P(gj))
    assert not s[12:].any()          # identity
    del r_as_fr


def test_group_law_vs_bigint(O):
    rnd = random.Random(3)
    g = O.generator()
    G = (fp_int(g[:6]), fp_int(g[6:]))
    for _ in range(20):
        a, b = rnd.randrange(1, 2**20), rnd.randrange(1, 2**20)
        pa = O.g1_mul(g, O.fr_from_int(a))
        pb = O.g1_mul(g, O.fr_from_int(b))
        s = O.g1j()
        O.lib().orc_g1j_add(O.P(s), O.P(pa), O.P(pb))
        # [a]G by Python double-and-add
        acc = None
        base = G
        e = a + b
        while e:
            if e & 1:
                acc = py_add(acc, base)
            base = py_add(base, base)
            e >>= 1
        assert affine_int(O, s) == acc
        assert O.lib().orc_g1a_is_on_curve(O.P(O.to_affine(s)))


def test_compress_roundtrip(O):
    pts = O.fixed_base(O.random_fr(50, 9))
    for p in pts:
        assert np.array_equal(O.decompress(O.compress(p)), p)
    assert O.compress(O.g1a())[0] == 0xc0


def test_srs_fixture(O):
    k, g, gl = O.load_srs_fixture(os.path.join(GOLDEN, "srs_2p14.g1c"))
    assert k == 14 and g.shape[0] == 1 << 14
    assert np.array_equal(g[0], O.generator())
    for arr in (g[:64], gl[:64]):
        for p in arr:
            assert O.lib().orc_g1a_is_on_curve(O.P(np.ascontiguousarray(p)))


@pytest.mark.skipif(not os.path.exists(REF_SRS), reason="reference tree not mounted")
def test_srs_fixture_matches_reference_file(O):
    k, g, gl = O.load_srs_fixture(os.path.join(GOLDEN, "srs_2p14.g1c"))
    k2, g2, gl2 = O.load_srs_file(REF_SRS)
    assert k == k2 and np.array_equal(g, g2) and np.array_equal(gl, gl2)


def test_chacha20_rfc7539_block(O):
    import ctypes
    key = (ctypes.c_uint32 * 8)(*[int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)])
    # RFC 7539 2.3.2: counter 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00.
    # The 64-bit-counter layout puts nonce word 0 in the counter's high half.
    counter = 1 | (0x09000000 << 32)
    nonce = (ctypes.c_uint32 * 2)(0x4a000000, 0)
    out = (ctypes.c_uint32 * 16)()
    O.lib().orc_chacha20_block(out, key, counter, nonce)
    want = [0xe4e7f110, 0x15593bd1, 0x1fdd0f50, 0xc47120a3, 0xc7f4d1c7, 0x0368c033, 0x9aaa2204, 0x4e6cd4c3,
            0x466482d2, 0x09aa9f07, 0x05d7c214, 0xa2028bd9, 0xd19c12b5, 0xb94e16de, 0xe883d0cb, 0x4e3c50a2]
    assert list(out) == want


# RFC 7539 appendix A.1, ChaCha20 block function test vectors #1 and #2:
# all-zero key and nonce, block counters 0 and 1 (serialized keystream)
RFC7539_A1_TV1 = bytes.fromhex(
    "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
    "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
RFC7539_A1_TV2 = bytes.fromhex(
    "9f07e7be5551387a98ba977c732d080dcb0f29a048e3656912c6533e32ee7aed"
    "29b721769ce64e43d57133b074d839d531ed1f28510afb45ace10a1f4b794d6f")


def _zero_seed_rng(O):
    import ctypes
    r = O.Rng()
    O.lib().orc_rng_from_seed(ctypes.byref(r), (ctypes.c_uint8 * 32)())
    return r


def test_chacha20_rng_stream_rfc7539_a1(O):
    """ChaCha20Rng::from_seed([0; 32]): its first 32 words are the RFC 7539
    A.1 keystream of blocks 0 and 1, little-endian; next_u64 pairs them low
    word first; fill_bytes serializes the words."""
    import ctypes
    ks = RFC7539_A1_TV1 + RFC7539_A1_TV2
    words = [int.from_bytes(ks[4 * i:4 * i + 4], "little") for i in range(32)]
    r = _zero_seed_rng(O)
    assert [O.lib().orc_rng_next_u32(ctypes.byref(r)) for _ in range(32)] == words
    r = _zero_seed_rng(O)
    assert [O.lib().orc_rng_next_u64(ctypes.byref(r)) for _ in range(16)] == \
        [words[2 * i] | words[2 * i + 1] << 32 for i in range(16)]
    rng = O.ChaCha(seed_bytes=bytes(32))
    assert rng.fill_bytes(64) == RFC7539_A1_TV1
    assert rng.fill_bytes(64) == RFC7539_A1_TV2


def test_chacha20_rng_u64_across_refill(O):
    """A u64 read at an odd word position: inside a refill it takes two
    consecutive words; at the last word of a 64-word refill it takes that
    word as its low half and the next refill's first word as its high half
    (rand_core BlockRng::next_u64), so the u64 stream is the word stream."""
    import ctypes
    ref = _zero_seed_rng(O)
    words = [O.lib().orc_rng_next_u32(ctypes.byref(ref)) for _ in range(200)]
    for skip in (1, 3, 61, 63, 65, 127):
        r = _zero_seed_rng(O)
        for _ in range(skip):
            O.lib().orc_rng_next_u32(ctypes.byref(r))
        got = [O.lib().orc_rng_next_u64(ctypes.byref(r)) for _ in range(4)]
        assert got == [words[skip + 2 * i] | words[skip + 2 * i + 1] << 32 for i in range(4)], skip
    # words 16..31 are RFC 7539 A.1 block 1: the 64-bit block counter counts through
    assert words[16] == int.from_bytes(RFC7539_A1_TV2[:4], "little")


def test_pcg32_demo_outputs(O):
    """The LCG step and XSH-RR output that seed_from_u64 expands a u64 with,
    against O'Neill's pcg32 demo: pcg32_srandom_r(42, 54) (inc = 54 * 2 + 1;
    state 0, step, add 42, step), then outputs taken from the state before
    each step."""
    lib = O.lib()
    inc = (54 << 1) | 1
    state = lib.orc_pcg32_advance(0, inc)
    state = lib.orc_pcg32_advance((state + 42) & (2**64 - 1), inc)
    out = []
    for _ in range(6):
        out.append(lib.orc_pcg32_output(state))
        state = lib.orc_pcg32_advance(state, inc)
    assert out == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_seed_from_u64_expansion(O):
    """seed_from_u64(s): eight PCG32 outputs of a state advanced before each
    output by the rand_core increment, serialized little-endian, as the
    ChaCha20 key."""
    lib = O.lib()
    inc = 11634580027462260723
    for s in (0, 1, 42, 2**64 - 1):
        st, seed = s, b""
        for _ in range(8):
            st = lib.orc_pcg32_advance(st, inc)
            seed += lib.orc_pcg32_output(st).to_bytes(4, "little")
        assert np.array_equal(O.ChaCha(seed=s).random_fr(5), O.ChaCha(seed_bytes=seed).random_fr(5)), s


@pytest.mark.skipif(not os.path.exists(REF_SRS), reason="reference tree not present")
def test_srs_file_rebuilt_from_fixture(O):
    """g, g_lagrange (uncompressed) and the G2 tail of the fixtures rebuild
    the reference's data file byte for byte."""
    k, g, gl = O.load_srs_fixture(os.path.join(GOLDEN, "srs_2p14.g1c"))
    with open(os.path.join(GOLDEN, "srs_2p14_g2.bin"), "rb") as f:
        tail = f.read()
    rebuilt = k.to_bytes(4, "little") + b"".join(O.to_uncompressed(p) for p in g) + \
        b"".join(O.to_uncompressed(p) for p in gl) + tail
    with open(REF_SRS, "rb") as f:
        assert f.read() == rebuilt
