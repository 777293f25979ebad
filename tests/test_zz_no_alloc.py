"""The no-allocation NTT entry point (midnight_hip_fr_device_vec_ntt_zeta_noalloc)
through the C ABI against the allocating entry point it is a variant of
(midnight_sppark_bls12_381_fr_device_vec_ntt_zeta): limb-identical results
on two calls in a row, the second smaller, forward and inverse, with and
without the zeta post-multiplication; sizes above its 2^20 scratch limit
are rejected.  The first call's size is checked against the oracle too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(H, x):
    return H.DeviceVec.upload(x), H.DeviceVec.upload(x)


@pytest.mark.parametrize("sizes", [(20, 14), (16, 11), (12, 5)])
def test_noalloc_ntt_matches_allocating_entry(H, O, sizes):
    one = O.fr_from_int(1)
    zeta = O.domain(5, 8)[0].fr("g_coset")
    zeta2 = O.fr_op("mul", zeta, zeta)
    scale = O.random_fr(1, 4000)[0]
    for call, log_n in enumerate(sizes):
        n = 1 << log_n
        x = O.random_fr(n, 4100 + 10 * log_n + call)
        for inverse, post in ((False, (one, one, one)), (True, (scale, zeta2, zeta))):
            print("noalloc call %d: 2^%d inverse=%s" % (call, log_n, inverse), flush=True)
            ref, got = _pair(H, x)
            ref.ntt_zeta(log_n, inverse, *post)
            got.ntt_zeta_noalloc(log_n, inverse, *post)
            want = ref.download()
            assert np.array_equal(got.download(), want), (call, log_n, inverse)
            if call == 0 and not inverse:
                assert np.array_equal(want, O.fft(x, log_n))
            ref.free()
            got.free()


def test_noalloc_ntt_rejects_sizes_above_limit(H, O):
    big = H.DeviceVec.upload(np.zeros((1 << 21, 4), dtype=np.uint64))
    one = O.fr_from_int(1)
    with pytest.raises(H.MidnightHipError):
        big.ntt_zeta_noalloc(21, False, one, one, one)
    with pytest.raises(H.MidnightHipError):
        big.ntt_zeta_noalloc(20, False, one, one, one)      # length mismatch
    big.free()
