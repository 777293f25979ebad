"""GPU parity: Fr NTT family through the C ABI vs the CPU oracle.

Bar: limb-identical (every value is a fully reduced Montgomery Fr).
Reference entry points: bls12_381_g1_msm.cu:6488 (fr_ntt), :3447 (ntt_zeta),
:3497 (divide_ntt_zeta), :2542 (coeff_to_extended_batch), :3281-3440
(device-vector variants); semantics vendor/midnight-proofs-0.8.1/src/poly/
domain.rs:325-443.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 8, 10, 11, 12, 13, 16, 17, 20])
def test_ntt_forward_inverse(H, O, log_n):
    n = 1 << log_n
    x = O.random_fr(n, 1000 + log_n)
    fwd = H.fr_ntt(x, log_n, inverse=False)
    assert np.array_equal(fwd, O.fft(x, log_n)), "forward NTT differs (log_n=%d)" % log_n
    inv = H.fr_ntt(x, log_n, inverse=True)
    assert np.array_equal(inv, O.ifft(x, log_n)), "inverse NTT differs (log_n=%d)" % log_n
    assert np.array_equal(H.fr_ntt(fwd, log_n, inverse=True), x)


@pytest.mark.parametrize("log_n", [4, 12, 15])
def test_ntt_extreme_values(H, O, log_n):
    """Inputs at the ends of the field (0, 1, r - 1, r - 2) and runs of
    them: the twiddle products (radix 2^29, twiddles scaled by 2^5) take
    their largest operands here."""
    n = 1 << log_n
    x = O.random_fr(n, 2000 + log_n)
    one = O.fr_from_int(1)
    r1 = O.fr_op("neg", one)
    r2 = O.fr_op("sub", r1, one)
    x[: n // 4] = r1
    x[n // 4: n // 4 + 7] = [0, 0, 0, 0]
    x[n // 2: n // 2 + 5] = one
    x[3 * n // 4:: 3] = r2
    for inv in (False, True):
        got = H.fr_ntt(x, log_n, inverse=inv)
        want = O.ifft(x, log_n) if inv else O.fft(x, log_n)
        assert np.array_equal(got, want), "log_n=%d inverse=%s" % (log_n, inv)


def test_ntt_2p22_roundtrip_and_spot(H, O):
    log_n = 22
    n = 1 << log_n
    x = O.random_fr(n, 77)
    fwd = H.fr_ntt(x, log_n)
    ref = O.fft(x, log_n)
    assert np.array_equal(fwd, ref)
    assert np.array_equal(H.fr_ntt(fwd, log_n, inverse=True), x)


def test_ntt_zeta_and_divide(H, O):
    d, t = O.domain(5, 12)     # j=5 -> extended_k = k + 2
    ek = d.extended_k
    N = 1 << ek
    v = O.random_fr(N, 5)
    scale = O.fr_op("mul", d.fr("extended_ifft_divisor"), O.fr_from_int(N))   # = 1, as the prover passes
    got = H.fr_ntt_zeta(v, ek, True, scale, d.fr("g_coset_inv"), d.fr("g_coset"))
    assert np.array_equal(got, O.extended_to_coeff(v, d))
    # divide_by_vanishing + extended_to_coeff
    got = H.fr_divide_ntt_zeta(v, ek, t, scale, d.fr("g_coset_inv"), d.fr("g_coset"))
    assert np.array_equal(got, O.extended_to_coeff(O.mul_repeating(v, t), d))


def test_coeff_to_extended_batch(H, O):
    d, _ = O.domain(5, 11)
    n, ek = 1 << d.k, d.extended_k
    polys = [O.random_fr(n, 40 + i) for i in range(3)]
    got = H.fr_coeff_to_extended_batch(np.concatenate(polys), n, 3, ek, d.fr("g_coset"), d.fr("g_coset_inv"))
    N = 1 << ek
    for i, p in enumerate(polys):
        assert np.array_equal(got[i * N:(i + 1) * N], O.coeff_to_extended(p, d))
    # resident variants
    dv = H.device_coeff_to_extended_batch(np.concatenate(polys), n, 3, ek, d.fr("g_coset"), d.fr("g_coset_inv"))
    assert np.array_equal(dv.download(), got)
    src = H.DeviceVec.upload(np.concatenate(polys))
    dv2 = H.device_coeff_to_extended_from_device_batch(src, n, 3, ek, d.fr("g_coset"), d.fr("g_coset_inv"))
    assert np.array_equal(dv2.download(), got)


def test_coeff_to_extended_short_input(H, O):
    # input shorter than the domain: zero padding beyond input_len
    d, _ = O.domain(3, 10)
    ek = d.extended_k
    N = 1 << ek
    short = O.random_fr(700, 9)
    got = H.fr_coeff_to_extended_batch(short, 700, 1, ek, d.fr("g_coset"), d.fr("g_coset_inv"))
    ref_in = np.zeros((N, 4), dtype=np.uint64)
    ref_in[:700] = O.mul_zeta(short, d.fr("g_coset"), d.fr("g_coset_inv"))
    assert np.array_equal(got, O.best_fft(ref_in, d.fr("extended_omega"), ek))


def test_device_vec_ntt_zeta_batch(H, O):
    log_n, count = 13, 3
    n = 1 << log_n
    polys = O.random_fr(n * count, 3)
    dv = H.DeviceVec.upload(polys)
    one = O.fr_from_int(1)
    dv.ntt_zeta_batch(log_n, count, True, one, one, one)
    got = dv.download()
    for i in range(count):
        assert np.array_equal(got[i * n:(i + 1) * n], O.ifft(polys[i * n:(i + 1) * n], log_n))


def test_ntt_multi_ragged(H, O):
    logs = [10, 12, 12, 11, 14, 10]
    offs, total = [], 0
    for lg in logs:
        offs.append(total)
        total += 1 << lg
    x = O.random_fr(total, 8)
    dv = H.DeviceVec.upload(x)
    dv.ntt_multi(offs, logs, inverse=True)
    got = dv.download()
    for o, lg in zip(offs, logs):
        assert np.array_equal(got[o:o + (1 << lg)], O.ifft(x[o:o + (1 << lg)], lg))


def test_vector_ops(H, O):
    n = 1000
    x = O.random_fr(n, 11)
    s = O.random_fr(1, 12)[0]
    a, b = O.random_fr(2, 13)
    assert np.array_equal(H.fr_mul_assign(x, s), O.mul_scalar(x, s))
    assert np.array_equal(H.fr_mul_assign_zeta(x, a, b), O.mul_zeta(x, a, b))
    f = O.random_fr(7, 14)
    assert np.array_equal(H.fr_mul_assign_repeating(x, f), O.mul_repeating(x, f))


def test_quotient_pieces(H, O):
    poly_len, pieces = 64, 4
    q = O.random_fr((poly_len - 1) * pieces, 21)
    blinds = O.random_fr(pieces - 1, 22)
    dv = H.DeviceVec.upload(q)
    got = dv.quotient_pieces(poly_len, pieces, blinds).download()
    # restate vanishing/prover.rs:163-182,304-318
    ref = np.zeros((poly_len * pieces, 4), dtype=np.uint64)
    for p in range(pieces):
        ref[p * poly_len:p * poly_len + poly_len - 1] = q[p * (poly_len - 1):(p + 1) * (poly_len - 1)]
    for i in range(pieces - 1):
        ref[i * poly_len + poly_len - 1] = blinds[i]
        ref[(i + 1) * poly_len] = O.fr_op("sub", ref[(i + 1) * poly_len], blinds[i])
    assert np.array_equal(got, ref)
