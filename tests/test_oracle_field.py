"""Pins the oracle's Fr / Fp arithmetic (CPU only).

Checks that do not restate the oracle's formula: the reference's own limb
KATs (vendor/midnight-curves-0.3.1/src/bls12_381/fq.rs:1147-1175,
fp.rs:1258-1290), its named constants (fq.rs:60-135, 476-505, 693-702),
Python big-integer arithmetic on random inputs, and closed forms
(Fermat inverse, root-of-unity orders, zeta^3 = 1).
"""
import random

import numpy as np

R_MOD = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001
P_MOD = 0x1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab


def limbs(v, n):
    return np.array([(v >> (64 * i)) & (2**64 - 1) for i in range(n)], dtype=np.uint64)


def as_int(a):
    return sum(int(x) << (64 * i) for i, x in enumerate(a))


def test_fr_mul_kat(O):
    a = np.array([0x6b7e9b8faeefc81a, 0xe30a8463f348ba42, 0xeff3cb67a8279c9c, 0x3d303651bd7c774d], dtype=np.uint64)
    b = np.array([0x13ae28e3bc35ebeb, 0xa10f4488075cae2c, 0x8160e95a853c3b5d, 0x5ae3f03b561a841d], dtype=np.uint64)
    want = [0x23717213ce710f71, 0xdbee1fe53a16e1af, 0xf565d3e1c2a48000, 0x4426507ee75df9d7]
    assert list(O.fr_op("mul", a, b)) == want


def test_fp_mul_kat(O):
    a = np.array([0xcc6200000020aa8a, 0x422800801dd8001a, 0x7f4f5e619041c62c, 0x8a55171ac70ed2ba,
                  0x3f69cc3a3d07d58b, 0xb972455fd09b8ef], dtype=np.uint64)
    b = np.array([0x329300000030ffcf, 0x633c00c02cc40028, 0xbef70d925862a942, 0x4f7fa2a82a963c17,
                  0xdf1eb2575b8bc051, 0x1162b680fb8e9566], dtype=np.uint64)
    want = [0x9dc4000001ebfe14, 0x2850078997b00193, 0xa8197f1abb4d7bf, 0xc0309573f4bfe871,
            0xf48d0923ffaf7620, 0x11d4b58c7a926e66]
    assert list(O.fp_op("mul", a, b)) == want


def test_fr_constants(O):
    one = O.fr()
    O.lib().orc_fr_one(O.P(one))
    assert list(one) == [0x00000001fffffffe, 0x5884b7fa00034802, 0x998c4fefecbc4ff5, 0x1824b159acc5056f]
    assert as_int(one) == (1 << 256) % R_MOD
    # R2 = from_u64(1) * R (fq.rs:99-108) via the conversion routine
    r2 = O.fr_from_int((1 << 256) % R_MOD)
    assert list(r2) == [0xc999e990f3f29c6d, 0x2b6cedcb87925c23, 0x05d314967254398f, 0x0748d9d99f59ff11]


def test_fr_random_vs_bigint(O):
    rnd = random.Random(5)
    for _ in range(200):
        x, y = rnd.randrange(R_MOD), rnd.randrange(R_MOD)
        a, b = O.fr_from_int(x), O.fr_from_int(y)
        assert as_int(a) == x * (1 << 256) % R_MOD         # Montgomery form
        assert O.fr_to_int(O.fr_op("mul", a, b)) == x * y % R_MOD
        assert O.fr_to_int(O.fr_op("add", a, b)) == (x + y) % R_MOD
        assert O.fr_to_int(O.fr_op("sub", a, b)) == (x - y) % R_MOD
        if x:
            assert O.fr_to_int(O.fr_op("inv", a)) == pow(x, R_MOD - 2, R_MOD)


def test_fp_random_vs_bigint(O):
    rnd = random.Random(6)
    R = 1 << 384
    for _ in range(100):
        x, y = rnd.randrange(P_MOD), rnd.randrange(P_MOD)
        a, b = limbs(x * R % P_MOD, 6), limbs(y * R % P_MOD, 6)
        assert as_int(O.fp_op("mul", a, b)) == x * y * R % P_MOD
        assert as_int(O.fp_op("add", a, b)) == (x + y) * R % P_MOD
        assert as_int(O.fp_op("sub", a, b)) == (x - y) * R % P_MOD
        if x:
            assert as_int(O.fp_op("inv", a)) == pow(x, P_MOD - 2, P_MOD) * R % P_MOD


def test_roots_of_unity(O):
    w = O.fr()
    O.lib().orc_fr_root_of_unity(O.P(w))
    wi = O.fr_to_int(w)
    assert pow(wi, 1 << 32, R_MOD) == 1 and pow(wi, 1 << 31, R_MOD) != 1
    z = O.fr()
    O.lib().orc_fr_zeta(O.P(z))
    zi = O.fr_to_int(z)
    assert zi != 1 and pow(zi, 3, R_MOD) == 1
    d = O.fr()
    O.lib().orc_fr_delta(O.P(d))
    t = (R_MOD - 1) >> 32
    assert pow(O.fr_to_int(d), t, R_MOD) == 1        # DELTA is a t-th root of unity
    for k in (1, 5, 14, 20):
        om = O.fr_to_int(O.omega(k))
        assert pow(om, 1 << k, R_MOD) == 1 and pow(om, 1 << (k - 1), R_MOD) == R_MOD - 1


def test_from_uniform_bytes_closed_form(O):
    rnd = random.Random(7)
    for _ in range(50):
        raw = bytes(rnd.randrange(256) for _ in range(64))
        out = O.fr()
        buf = np.frombuffer(raw, dtype=np.uint8).copy()
        O.lib().orc_fr_from_uniform_bytes(O.P(out), O.B(buf))
        assert O.fr_to_int(out) == int.from_bytes(raw, "little") % R_MOD


def test_random_fr_is_reduced_and_seeded(O):
    a = O.random_fr(1000, 42)
    b = O.random_fr(1000, 42)
    assert np.array_equal(a, b)
    assert all(as_int(x) < R_MOD for x in a)
    assert not np.array_equal(a, O.random_fr(1000, 43))
