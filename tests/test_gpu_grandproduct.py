"""GPU parity: grand products through the C ABI vs the CPU oracle.

Bar: limb-identical.  Reference entry points: bls12_381_g1_msm.cu:4303
(fr_shifted_prefix_product), :4342 (fr_permutation_products), :5738
(trace_context_permutation); semantics plonk/permutation/prover.rs:36-43,
355-515.  The oracle's products are pinned by tests/test_oracle_quotient.py
(a satisfying circuit's last product closes to one).
"""
import numpy as np
import pytest

import synth_circuit as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 2047, 2048, 2049, 5000, 1 << 16, (1 << 20) + 3])
def test_shifted_prefix_product(H, O, n):
    v = O.random_fr(n, 100 + n % 97)
    init = O.random_fr(1, 7)[0]
    print("shifted prefix product n=%d" % n, flush=True)
    assert np.array_equal(H.fr_shifted_prefix_product(v, init), O.shifted_prefix_product(v, init))


def _delta_powers(O, count):
    d = O.fr()
    O.lib().orc_fr_delta(O.P(d))
    out = [O.fr_from_int(1)]
    for _ in range(count - 1):
        out.append(O.fr_op("mul", out[-1], d))
    return np.stack(out)


def test_permutation_products_satisfied_circuit(H, O):
    c = S.Circuit(9, seed=3)
    cols = S.mont([x for col in c.cols for x in col])
    sig = S.mont([x for col in c.sigma for x in col])
    kw = dict(beta=S.mont([c.beta])[0], gamma=S.mont([c.gamma])[0], omega=S.mont([c.omega])[0],
              delta_powers=S.mont(c.delta_pows), blinds=S.mont(c.blinds), blinding_factors=S.BLINDING)
    dv = H.fr_permutation_products(cols, sig, c.n, S.CHUNK, **kw)
    got = dv.download()
    assert np.array_equal(got, S.mont([x for z in c.z for x in z]))
    assert S.ints(got[2 * c.n + c.u][None, :])[0] == 1
    dv.free()


@pytest.mark.parametrize("log_n,ncols,chunk,bf", [(12, 8, 3, 5), (16, 8, 3, 5), (13, 7, 2, 0), (20, 8, 3, 5)])
def test_permutation_products_random(H, O, log_n, ncols, chunk, bf):
    n = 1 << log_n
    sets = -(-ncols // chunk)
    vals = O.random_fr(2 * ncols * n + sets * bf + 4, 500 + log_n)
    cols, sig = vals[:ncols * n], vals[ncols * n:2 * ncols * n]
    blinds = vals[2 * ncols * n:2 * ncols * n + sets * bf]
    beta, gamma = vals[-1], vals[-2]
    omega = O.omega(log_n)
    dp = _delta_powers(O, ncols)
    print("permutation products n=2^%d cols=%d chunk=%d" % (log_n, ncols, chunk), flush=True)
    dv = H.fr_permutation_products(cols, sig, n, chunk, beta=beta, gamma=gamma, omega=omega, delta_powers=dp,
                                   blinds=blinds, blinding_factors=bf)
    want = O.permutation_products(cols, sig, n, chunk, beta, gamma, omega, dp, blinds, bf)
    assert np.array_equal(dv.download(), want)
    dv.free()


def test_permutation_products_zero_denominator(H, O):
    # a row whose denominator vanishes: the fraction there is zero
    n = 4096
    cols = O.random_fr(3 * n, 601)
    sig = O.random_fr(3 * n, 602)
    beta, gamma = O.random_fr(2, 603)
    row = 1234
    # v + beta * sigma + gamma = 0 for column 0 at `row`
    bs = O.fr_op("mul", beta, sig[row])
    cols[row] = O.fr_op("neg", O.fr_op("add", bs, gamma))
    dp = _delta_powers(O, 3)
    dv = H.fr_permutation_products(cols, sig, n, 3, beta=beta, gamma=gamma, omega=O.omega(12), delta_powers=dp,
                                   blinds=O.random_fr(5, 604), blinding_factors=5)
    want = O.permutation_products(cols, sig, n, 3, beta, gamma, O.omega(12), dp, O.random_fr(5, 604), 5)
    got = dv.download()
    assert np.array_equal(got, want)
    assert not got[row + 1].any()
    dv.free()


def test_trace_context_permutation(H, O):
    """Products from a numerator context's Lagrange fixed / sigma columns and
    resident advice (the resident prover path)."""
    c = S.Circuit(8, seed=17)
    n = c.n
    ext = 4 * n
    filler = O.random_fr(ext * 14, 700)
    ctx = H.NumeratorContext(
        ext, n, fixed=filler[:3 * ext], fixed_lagrange=S.mont([x for col in c.cols[5:] for x in col]),
        l0=filler[3 * ext:4 * ext], l_last=filler[4 * ext:5 * ext], l_active=filler[5 * ext:6 * ext],
        permutation_columns=S.PERM_COLUMNS, permutation_pk_cosets=filler[6 * ext:14 * ext],
        permutation_pk_lagrange=S.mont([x for col in c.sigma for x in col]),
        custom_graph=c.custom_graph(), omega=O.omega(12), trace_omega=S.mont([c.omega])[0],
        g_coset=O.fr_from_int(1), delta_const=O.fr_from_int(1), rot_scale=4, last_rotation=-6, chunk_len=S.CHUNK)
    adv = H.DeviceVec.upload(S.mont([x for col in c.cols[:5] for x in col]))
    dv = ctx.trace_permutation(adv, 5, None, 0, beta=S.mont([c.beta])[0], gamma=S.mont([c.gamma])[0],
                               delta_powers=S.mont(c.delta_pows), blinds=S.mont(c.blinds),
                               blinding_factors=S.BLINDING)
    assert np.array_equal(dv.download(), S.mont([x for z in c.z for x in z]))
    with pytest.raises(H.MidnightHipError):
        ctx.trace_permutation(adv, 4, None, 0, beta=S.mont([c.beta])[0], gamma=S.mont([c.gamma])[0],
                              delta_powers=S.mont(c.delta_pows), blinds=S.mont(c.blinds),
                              blinding_factors=S.BLINDING)
    dv.free()
    adv.free()
    ctx.free()
