"""Pins the oracle's NTT, domain and MSM (CPU only).

- NTT: the DFT definition evaluated directly (Python big integers) on small
  sizes; iNTT(NTT(x)) = x; NTT of a monomial x^j is (omega^(ij))_i.
- Domain (vendor/midnight-proofs-0.8.1/src/poly/domain.rs:54-159): the
  extended degree rule, t_evaluations = 1 / ((zeta * w_ext^i)^n - 1), and
  extended_to_coeff(coeff_to_extended(p)) = p.
- MSM: Pippenger vs independent double-and-add, the closed form
  MSM([t_i]G, s) = [sum s_i t_i]G, and the reference SRS data file identity
  MSM(g, c) == MSM(g_lagrange, NTT(c)) (g_lagrange = iFFT(g),
  vendor/midnight-proofs-0.8.1/src/utils/arithmetic.rs:34-50).
- Booth digits (midnight-curves msm.rs:13-50): sum_w d_w 2^(c w) = scalar.
"""
import os

import numpy as np
import pytest

R_MOD = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("log_n", [0, 1, 2, 3, 5])
def test_ntt_is_the_dft(O, log_n):
    n = 1 << log_n
    x = O.random_fr(n, 10 + log_n)
    xi = [O.fr_to_int(v) for v in x]
    w = O.fr_to_int(O.omega(log_n))
    want = [sum(xi[j] * pow(w, i * j, R_MOD) for j in range(n)) % R_MOD for i in range(n)]
    got = O.fft(x, log_n)
    assert [O.fr_to_int(v) for v in got] == want


def test_ntt_roundtrip_and_monomial(O):
    log_n = 10
    n = 1 << log_n
    x = O.random_fr(n, 1)
    assert np.array_equal(O.ifft(O.fft(x, log_n), log_n), x)
    j = 37
    mono = np.zeros((n, 4), dtype=np.uint64)
    mono[j] = O.fr_from_int(1)
    got = O.fft(mono, log_n)
    w = O.fr_to_int(O.omega(log_n))
    for i in (0, 1, 5, 511, 1023):
        assert O.fr_to_int(got[i]) == pow(w, i * j, R_MOD)


def test_domain_constants(O):
    for j, k, ek in ((5, 10, 12), (4, 10, 12), (3, 10, 11), (2, 10, 10)):
        d, t = O.domain(j, k)
        assert d.extended_k == ek
        n = 1 << k
        z = O.fr_to_int(d.fr("g_coset"))
        wext = O.fr_to_int(d.fr("extended_omega"))
        assert pow(z, 3, R_MOD) == 1 and O.fr_to_int(d.fr("g_coset_inv")) == z * z % R_MOD
        for i in range(len(t)):
            v = (pow(z * pow(wext, i, R_MOD), n, R_MOD) - 1) % R_MOD
            assert O.fr_to_int(t[i]) == pow(v, R_MOD - 2, R_MOD)


def test_coset_roundtrip(O):
    d, _ = O.domain(5, 9)
    p = O.random_fr(1 << 9, 4)
    ext = O.coeff_to_extended(p, d)
    back = O.extended_to_coeff(ext, d)
    assert np.array_equal(back[:1 << 9], p)
    assert not back[1 << 9:].any()
    # evaluation at the first coset point: p(zeta)
    z = O.fr_to_int(d.fr("g_coset"))
    pi = [O.fr_to_int(c) for c in p]
    assert O.fr_to_int(ext[0]) == sum(c * pow(z, i, R_MOD) for i, c in enumerate(pi)) % R_MOD


def test_booth_digits_recompose(O):
    for c in (3, 8, 13, 14):
        for s in O.random_fr(20, c):
            can = O.fr_to_int(s)
            raw = np.frombuffer(can.to_bytes(32, "little"), dtype=np.uint8).copy()
            nw = 255 // c + 1
            total = sum(O.lib().orc_booth_index(w, c, O.B(raw)) * (1 << (c * w)) for w in range(nw))
            assert total == can


@pytest.mark.parametrize("n", [1, 5, 31, 32, 200, 3000])
def test_pippenger_vs_naive_and_closed_form(O, n):
    ts = O.random_fr(n, 100 + n)
    pts = O.fixed_base(ts)
    sc = O.random_fr(n, 200 + n)
    a = O.msm(pts, sc)
    if n <= 200:
        assert O.g1_eq(a, O.msm(pts, sc, naive=True))
    assert O.g1_eq(a, O.g1_mul(O.generator(), O.inner_product(sc, ts)))


def test_fixed_base_matches_scalar_mul(O):
    ts = O.random_fr(8, 77)
    pts = O.fixed_base(ts)
    for t, p in zip(ts, pts):
        assert np.array_equal(O.to_affine(O.g1_mul(O.generator(), t)), p)


def test_srs_identity_on_reference_data(O):
    k, g, gl = O.load_srs_fixture(os.path.join(GOLDEN, "srs_2p14.g1c"))
    c = O.random_fr(1 << k, 99)
    assert O.g1_eq(O.msm(g, c), O.msm(gl, O.fft(c, k)))
    # a deliberately wrong NTT direction must fail the identity
    assert not O.g1_eq(O.msm(g, c), O.msm(gl, O.ifft(c, k)))


def test_fp_inverse_vartime(O):
    """The binary-EEA inverse the batch-affine schedule uses equals the
    Fermat inverse, and a * a^-1 = 1."""
    one = np.zeros(6, dtype=np.uint64)
    O.lib().orc_fp_one(O.P(one))
    pts = O.fixed_base(O.random_fr(40, 555))
    for p in pts:
        x = np.ascontiguousarray(p[:6])
        assert np.array_equal(O.fp_op("inv_vartime", x), O.fp_op("inv", x))
        assert np.array_equal(O.fp_op("mul", x, O.fp_op("inv_vartime", x)), one)
    assert np.array_equal(O.fp_op("inv_vartime", one), one)


@pytest.mark.parametrize("n", [8200, 20000])
def test_msm_best_batch_affine(O, n):
    """msm_best's batch-affine schedule (c = ceil(ln n) >= 10) against the
    closed form, with doublings (a point twice under equal scalars), P + (-P)
    cancellations (a point and its negation under one scalar) and heavy
    buckets (runs of one scalar) in the mix."""
    ts = O.random_fr(n, 300 + n)
    sc = O.random_fr(n, 400 + n)
    r1 = O.fr_op("neg", O.fr_from_int(1))
    ts[100:160] = ts[0]                     # one point many times, one scalar: doublings
    sc[100:160] = sc[0]
    ts[200:260] = O.fr_op("neg", ts[1])     # -P beside P under one scalar: cancellations
    sc[200:260] = sc[1]
    sc[500:1500] = sc[2]                    # a heavy bucket in every window
    sc[2000] = 0
    sc[2001] = r1
    pts = O.fixed_base(ts)
    got = O.msm(pts, sc, best=True)
    assert O.g1_eq(got, O.g1_mul(O.generator(), O.inner_product(sc, ts)))
    assert O.g1_eq(got, O.msm(pts, sc))
