"""GPU parity: the plain KZG commitment workload (2^16 Lagrange values
drawn from a ChaCha20 seed with Fq::random, inverse NTT, MSM against the
SRS [tau^i]G with tau from the same seed) equals the closed form
[p(tau)]G, p the interpolating polynomial (commit_lagrange,
vendor/midnight-proofs-0.8.1/src/poly/kzg/mod.rs:1072-1770)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log_n", [10, 16])
def test_kzg_commit_closed_form(H, O, log_n):
    n = 1 << log_n
    tau = O.random_fr(1, 1)[0]
    evals = O.random_fr(n + 1, 1)[1:]          # the seed's stream after tau
    powers = H.fr_shifted_prefix_product(np.tile(tau, (n, 1)), O.fr_from_int(1))
    assert np.array_equal(powers[2], O.fr_op("mul", tau, tau))
    pw = H.DeviceVec.upload(powers)
    ctx = H.MsmContext.from_multiples(pw)
    dv = H.DeviceVec.upload(evals)
    one = O.fr_from_int(1)
    dv.ntt_zeta(log_n, True, one, one, one)
    coeffs = dv.download()
    assert np.array_equal(coeffs, O.ifft(evals, log_n))
    com = ctx.device_batch(dv, 0, n, 1)[0]
    want = O.g1_mul(O.generator(), O.inner_product(coeffs, powers))
    assert O.g1_eq(com, want)
    # the SRS points are [tau^i]G
    assert np.array_equal(ctx.points(4), O.srs(tau, 4))
    ctx.free()
    dv.free()
    pw.free()
