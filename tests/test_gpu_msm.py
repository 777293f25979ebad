"""GPU parity: G1 MSM through the C ABI vs the CPU oracle.

Bar: equal group elements (the Jacobian outputs are compared after
normalisation, as the reference's callers do; the affine result is unique).
Reference entry points: bls12_381_g1_msm.cu:6164-6486.  Test patterns follow
midnight-sppark-msm/src/lib.rs:5277-5379 (points (i+1)G, scalars with 1/7
zeros) and the reference's SRS data file (MSM(g, c) == MSM(g_lagrange,
NTT(c)), tests/golden/srs_2p14.g1c).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def points_for(O, ts):
    return O.fixed_base(ts)


def closed_form(O, ts, scalars):
    """MSM([t_i]G, s) = [sum s_i t_i]G"""
    return O.g1_mul(O.generator(), O.inner_product(scalars, ts))


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 100, 1000, 4097])
def test_msm_small_vs_naive(H, O, n):
    ts = O.random_fr(n, 500 + n)
    pts = points_for(O, ts)
    sc = O.random_fr(n, 600 + n)
    got = H.g1_msm(pts, sc)
    ref = O.msm(pts, sc, naive=n <= 100)
    assert O.g1_eq(got, ref), "n=%d" % n
    assert O.g1_eq(got, closed_form(O, ts, sc))


def test_msm_reference_pattern(H, O):
    # lib.rs:5277-5379: points (i+1)G, every 7th scalar zero
    n = 4096
    ts = np.stack([O.fr_from_int(i + 1) for i in range(n)])
    pts = points_for(O, ts)
    sc = O.random_fr(n, 9)
    sc[::7] = 0
    got = H.g1_msm(pts, sc)
    assert O.g1_eq(got, closed_form(O, ts, sc))


def test_msm_edge_scalars_and_points(H, O):
    n = 300
    ts = O.random_fr(n, 21)
    pts = points_for(O, ts)
    sc = O.random_fr(n, 22)
    r_minus_1 = O.fr_op("neg", O.fr_from_int(1))
    sc[0] = 0                       # zero scalar
    sc[1] = O.fr_from_int(1)        # one
    sc[2] = r_minus_1               # largest scalar
    sc[3] = r_minus_1
    pts[4] = 0                      # identity point (encoded as 0,0)
    pts[5] = pts[6]                 # duplicate point, same scalar -> doubling in a bucket
    sc[5] = sc[6]
    ts[4] = 0
    ts[5] = ts[6]
    pts[7] = pts[8].copy()          # P and -P with equal scalars -> cancel
    neg_y = O.fp_op("sub", np.zeros(6, dtype=np.uint64), pts[8][6:])
    pts[7][6:] = neg_y
    ts[7] = O.fr_op("neg", ts[8])
    sc[7] = sc[8]
    got = H.g1_msm(pts, sc)
    assert O.g1_eq(got, closed_form(O, ts, sc))
    assert O.g1_eq(got, O.msm(pts, sc, naive=True))


def test_msm_all_zero_and_all_equal(H, O):
    n = 2000
    ts = O.random_fr(n, 31)
    pts = points_for(O, ts)
    z = np.zeros((n, 4), dtype=np.uint64)
    got = H.g1_msm(pts, z)
    assert not got[12:].any()          # identity: z = 0
    # every scalar equal: one bucket per window holds every point
    one = np.tile(O.fr_from_int(1), (n, 1))
    assert O.g1_eq(H.g1_msm(pts, one), closed_form(O, ts, one))
    big = np.tile(O.random_fr(1, 32)[0], (n, 1))
    assert O.g1_eq(H.g1_msm(pts, big), closed_form(O, ts, big))


def test_msm_srs_fixture_identity(H, O):
    """Reference data: MSM(g, c) == MSM(g_lagrange, NTT(c)) at k=14."""
    k, g, gl = O.load_srs_fixture(os.path.join(GOLDEN, "srs_2p14.g1c"))
    n = 1 << k
    c = O.random_fr(n, 1234)
    e = H.fr_ntt(c, k)             # GPU NTT
    a = H.g1_msm(g, c)             # GPU MSM, monomial basis
    b = H.g1_msm(gl, e)            # GPU MSM, Lagrange basis
    assert O.g1_eq(a, b)
    assert O.g1_eq(a, O.msm(g, c))


def test_msm_bucket_tree_capped_levels(H, O):
    """Buckets deeper than the launched tree levels: every scalar of a batch
    equal (each window's digit bucket holds every point), run with the tree
    capped at 0 and 1 levels and uncapped; the last reduction pass sums what
    the levels leave, and every result is the closed form."""
    n, m = 4096, 3
    ts = O.random_fr(n, 1601)
    dv_t = H.DeviceVec.upload(ts)
    ctx = H.MsmContext.from_multiples(dv_t)
    dv_t.free()
    sc = O.random_fr(n * m, 1602)
    sc[:n] = sc[0]
    sc[2 * n:] = sc[2 * n]
    dv = H.DeviceVec.upload(sc)
    want = [closed_form(O, ts, sc[i * n:(i + 1) * n]) for i in range(m)]
    try:
        for cap in (0, 1, 0xffffffff):
            H.lib().midnight_hip_test_msm_tree_levels(cap)
            got = ctx.device_batch(dv, 0, n, m)
            for i in range(m):
                assert O.g1_eq(got[i], want[i]), (cap, i)
    finally:
        H.lib().midnight_hip_test_msm_tree_levels(0xffffffff)
    dv.free()
    ctx.free()


def test_msm_context_batch_variants(H, O):
    n, m = 3000, 4
    ts = O.random_fr(n, 41)
    pts = points_for(O, ts)
    sc = O.random_fr(n * m, 42)
    ctx = H.MsmContext(pts)
    out = ctx.batch(sc, n, m)
    for i in range(m):
        assert O.g1_eq(out[i], closed_form(O, ts, sc[i * n:(i + 1) * n]))
    # resident scalars
    dv = H.DeviceVec.upload(np.concatenate([O.random_fr(5, 43), sc]))
    out2 = ctx.device_batch(dv, 5, n, m)
    for i in range(m):
        assert O.g1_eq(out2[i], out[i])
    # ragged: (scalar_offset, point_offset, len, nonzero_len)
    slices = [(5, 0, n, n), (5 + n, 100, 500, 500), (5 + 2 * n, 7, 0, 0), (5 + 3 * n, 2000, 1000, 10)]
    out3 = ctx.device_ragged(dv, slices)
    allsc = dv.download()
    for (so, po, ln, nz), res in zip(slices, out3):
        ref = closed_form(O, ts[po:po + ln], allsc[so:so + ln]) if ln else None
        if ln == 0:
            assert not res[12:].any()      # identity: z = 0
        else:
            assert O.g1_eq(res, ref)
    # scrambled: returns MSM(P, s + mask)
    mask = O.random_fr(n, 44)
    mdv = H.DeviceVec.upload(mask)
    out4 = ctx.device_scrambled_batch(dv, 5, n, 2, mdv)
    for i in range(2):
        s_plus = np.stack([O.fr_op("add", sc[i * n + j], mask[j]) for j in range(n)])
        assert O.g1_eq(out4[i], closed_form(O, ts, s_plus))
    # batched ragged sizes over one resident vector
    sizes = [n, 1024, 17, 2999]
    offs = [5, 5 + n, 5 + 2 * n, 5]
    out5 = ctx.device_multi(dv, offs, sizes)
    for o, s, res in zip(offs, sizes, out5):
        assert O.g1_eq(res, closed_form(O, ts[:s], allsc[o:o + s]))
    ctx.free()


def test_msm_batch_entry(H, O):
    n, m = 1500, 3
    ts = O.random_fr(n, 51)
    pts = points_for(O, ts)
    sc = O.random_fr(n * m, 52)
    out = H.g1_msm_batch(pts, sc, m)
    for i in range(m):
        assert O.g1_eq(out[i], closed_form(O, ts, sc[i * n:(i + 1) * n]))


def test_msm_batch_many_small_msms(H, O):
    """1100 one-shot MSMs of 64 shared points: plain 8-bit windows give
    1100 x 32 x 128 = 4.5M bucket keys, more than the two-pass placement
    covers (4096 keys per coarse bin x 1024 bins), so the direct placement
    pass runs."""
    n, m = 64, 1100
    ts = O.random_fr(n, 53)
    pts = points_for(O, ts)
    sc = O.random_fr(n * m, 54)
    out = H.g1_msm_batch(pts, sc, m)
    for i in range(0, m, 7):
        assert O.g1_eq(out[i], closed_form(O, ts, sc[i * n:(i + 1) * n]))
    assert O.g1_eq(out[m - 1], closed_form(O, ts, sc[(m - 1) * n:]))


@pytest.mark.parametrize("log_n", [16, 18])
def test_msm_medium_closed_form(H, O, log_n):
    n = 1 << log_n
    ts = O.random_fr(n, 60 + log_n)
    pts = points_for(O, ts)
    sc = O.random_fr(n, 70 + log_n)
    assert O.g1_eq(H.g1_msm(pts, sc), closed_form(O, ts, sc))


def test_msm_skewed_mix(H, O):
    # the skewed scalar mix of the 2^24 workload, at 2^16
    n = 1 << 16
    ts = O.random_fr(n, 80)
    pts = points_for(O, ts)
    sc = O.random_fr(n, 81)
    rng = np.random.default_rng(82)
    u = rng.random(n)
    small = u < 0.7
    zo = (u >= 0.7) & (u < 0.9)
    for i in np.nonzero(small)[0]:
        sc[i] = O.fr_from_int(int(rng.integers(0, 1 << 16)))
    for i in np.nonzero(zo)[0]:
        sc[i] = O.fr_from_int(int(rng.integers(0, 2)))
    assert O.g1_eq(H.g1_msm(pts, sc), closed_form(O, ts, sc))


def test_msm_errors(H, O):
    # null context (bls12_381_g1_msm.cu:6295-6297)
    with pytest.raises(H.MidnightHipError):
        H.check(H.lib().midnight_sppark_bls12_381_g1_msm_context_batch(None, None, None, 1, 1))
    bad = H.RawMsmConfig(14, 100000, 32768, 9, 3, 2, 1, 4, 0, 1)   # window_bits 9 is invalid
    with pytest.raises(H.MidnightHipError):
        H.MsmContext(np.zeros((4, 12), dtype=np.uint64), bad)


def test_generator_multiples_and_montgomery(H, O):
    n = 5000
    ts = O.random_fr(n, 91)
    # canonical -> Montgomery on the device
    can = np.stack([np.array([(O.fr_to_int(t) >> (64 * i)) & (2**64 - 1) for i in range(4)], dtype=np.uint64)
                    for t in ts])
    dv = H.DeviceVec.upload(can)
    H.to_montgomery(dv)
    assert np.array_equal(dv.download(), ts)
    ctx = H.MsmContext.from_multiples(dv)
    assert np.array_equal(ctx.points(), O.fixed_base(ts))
    sc = O.random_fr(n, 92)
    assert O.g1_eq(ctx.batch(sc, n, 1)[0], O.g1_mul(O.generator(), O.inner_product(sc, ts)))
    entries, window = ctx.last_stats()
    assert window >= 8 and 0 < entries <= n * (-(-256 // window))
    ctx.free()


@pytest.mark.parametrize("log_n", [10, 14, 17, 21])
def test_msm_precomputed_context_modes(H, O, log_n):
    """Resident contexts use per-window point copies; compare against the
    one-shot (plain window) path and the closed form, with skewed and
    uniform scalars, several MSMs per call.  2^21 points select a 17-bit
    window (wider than any one-shot window)."""
    n = 1 << log_n
    ts = O.random_fr(n, 300 + log_n)
    dv_t = H.DeviceVec.upload(ts)
    ctx = H.MsmContext.from_multiples(dv_t)
    sc = O.random_fr(3 * n, 310 + log_n)
    sc[n:2 * n:3] = 0
    sc[2 * n:] = np.tile(O.fr_from_int(1), (n, 1))        # one bucket holds every point
    out = ctx.batch(sc, n, 3)
    for i in range(3):
        assert O.g1_eq(out[i], O.g1_mul(O.generator(), O.inner_product(sc[i * n:(i + 1) * n], ts)))
    pts = ctx.points()
    assert O.g1_eq(H.g1_msm(pts, sc[:n]), out[0])
    ctx.free()


def test_msm_precomputed_repeats_and_cancellations(H, O):
    """Precomputed-context accumulation (loose coordinates in [0, 2p)) on
    buckets that repeat one point (doubling inside a chunk: the difference
    of equal x coordinates may come out as p rather than 0) and that
    alternate Q and -Q (the running sum returns to the identity mid-chunk
    and restarts)."""
    n = 1 << 12
    ts = O.random_fr(n, 330)
    sc = O.random_fr(n, 331)
    t0, s0 = ts[0].copy(), sc[0].copy()
    ts[:1024] = t0                                      # one point 1024 times, one scalar
    sc[:1024] = s0
    t1, s1 = ts[1024].copy(), sc[1024].copy()
    neg_t1 = O.fr_op("neg", t1)
    ts[1024:2048:2] = neg_t1                            # -Q, Q, -Q, Q, ... with equal scalars
    ts[1025:2048:2] = t1
    sc[1024:2048] = s1
    ctx = H.MsmContext.from_multiples(H.DeviceVec.upload(ts))
    out = ctx.batch(sc, n, 1)
    assert O.g1_eq(out[0], O.g1_mul(O.generator(), O.inner_product(sc, ts)))
    ctx.free()


def _skewed_montgomery(O, sc, seed):
    """The 2^24 workload's skewed mix, in place on Montgomery scalars:
    70% in [0, 2^16), 20% in {0, 1}, 10% left uniform."""
    rng = np.random.default_rng(seed)
    table = np.stack([O.fr_from_int(v) for v in range(1 << 16)])
    u = rng.random(sc.shape[0])
    small = u < 0.7
    zo = (u >= 0.7) & (u < 0.9)
    sc[small] = table[rng.integers(0, 1 << 16, size=int(small.sum()))]
    sc[zo] = table[rng.integers(0, 2, size=int(zo.sum()))]
    return sc


@pytest.mark.slow
def test_msm_2p24_uniform_and_skewed(H, O):
    """2^24 resident points [t_i]G, uniform and skewed scalars in one call:
    each result equals [<s, t>]G."""
    n = 1 << 24
    print("MSM 2^24: generating inputs", flush=True)
    ts = O.random_fr(n, 2401)
    dv_t = H.DeviceVec.upload(ts)
    ctx = H.MsmContext.from_multiples(dv_t)
    dv_t.free()
    uni = O.random_fr(n, 2402)
    skw = _skewed_montgomery(O, O.random_fr(n, 2403), 2404)
    print("MSM 2^24: device batch", flush=True)
    dv = H.DeviceVec.upload(np.concatenate([uni, skw]))
    out = ctx.device_batch(dv, 0, n, 2)
    for i, sc in enumerate((uni, skw)):
        assert O.g1_eq(out[i], O.g1_mul(O.generator(), O.inner_product(sc, ts))), i
    dv.free()
    ctx.free()


def test_msm_short_on_large_precomputed_context(H, O):
    """MSMs much shorter than a precomputed context's bucket set take plain
    windows over the context's points (a small circuit over a large SRS);
    mixed lengths in one ragged batch, each against the closed form."""
    n = 1 << 16
    ts = O.random_fr(n, 990)
    dv_t = H.DeviceVec.upload(ts)
    ctx = H.MsmContext.from_multiples(dv_t)
    sizes = [1, 100, 1 << 10, 1 << 12, 1 << 14, 1 << 16]
    sc = O.random_fr(sum(sizes), 991)
    dv = H.DeviceVec.upload(sc)
    offs = np.cumsum([0] + sizes[:-1]).tolist()
    out = ctx.device_multi(dv, offs, sizes)
    for i, (o, m) in enumerate(zip(offs, sizes)):
        want = O.g1_mul(O.generator(), O.inner_product(sc[o:o + m], ts[:m]))
        assert O.g1_eq(out[i], want), m
    for m in (1 << 12, 1 << 15):
        got = ctx.device_batch(dv, 0, m, 1)[0]
        assert O.g1_eq(got, O.g1_mul(O.generator(), O.inner_product(sc[:m], ts[:m]))), m
    ctx.free()
    dv.free()
    dv_t.free()


def test_msm_small_table_on_wide_context(H, O):
    """A context whose main table is wider than 13 bits (2^18 points) also
    holds a 13-bit table over its first 2^(c-1) points: batches of short
    MSMs within it (ragged, uniform and repeated scalars) take that table,
    each against the closed form, and either side of the switch to the
    main table."""
    n = 1 << 18
    ts = O.random_fr(n, 1210)
    dv_t = H.DeviceVec.upload(ts)
    ctx = H.MsmContext.from_multiples(dv_t)
    w, _, _ = ctx.tuning()
    assert w > 13
    m = (1 << 12) - 1                                   # below every main window's switch
    sizes = [1, 100, 1000, m, m]
    sc = O.random_fr(sum(sizes), 1211)
    sc[sizes[0] + sizes[1]:sum(sizes[:3])] = sc[sizes[0] + sizes[1]]     # one scalar 1000 times
    dv = H.DeviceVec.upload(sc)
    offs = np.cumsum([0] + sizes[:-1]).tolist()
    out = ctx.device_multi(dv, offs, sizes)
    assert ctx.last_stats()[1] == 13
    for i, (o, m) in enumerate(zip(offs, sizes)):
        assert O.g1_eq(out[i], O.g1_mul(O.generator(), O.inner_product(sc[o:o + m], ts[:m]))), m
    got = ctx.device_batch(dv, sum(sizes[:3]), m, 2)
    assert ctx.last_stats()[1] == 13
    for i in range(2):
        o = sum(sizes[:3]) + i * m
        assert O.g1_eq(got[i], O.g1_mul(O.generator(), O.inner_product(sc[o:o + m], ts[:m]))), i
    dv.free()
    # either side of the switch to the main table at 2^(w-1) points
    big = 1 << (w - 1)
    sc2 = O.random_fr(big, 1212)
    dv2 = H.DeviceVec.upload(sc2)
    for m, window in ((big - 1, 13), (big, w)):
        got2 = ctx.device_batch(dv2, 0, m, 1)[0]
        assert ctx.last_stats()[1] == window, m
        assert O.g1_eq(got2, O.g1_mul(O.generator(), O.inner_product(sc2[:m], ts[:m]))), m
    dv2.free()
    ctx.free()
    dv_t.free()


def _reference_points(O, n):
    """points (i + 1) G, as the reference's ordering tests use"""
    ts = np.stack([O.fr_from_int(i + 1) for i in range(n)])
    return ts, points_for(O, ts)


@pytest.mark.parametrize("ragged,sparse", [(False, False), (True, False), (True, True)])
def test_device_msm_waits_for_delayed_input_producer(H, O, ragged, sparse):
    """lib.rs:5110-5199: the scalars' producer (a scaling by 2) is held back
    100 ms on the lane's stream and has not run when the MSM is issued; the
    MSM must see the doubled scalars."""
    n, m = 4096, 3
    ts, pts = _reference_points(O, n)
    ctx = H.MsmContext(pts)
    sc = np.stack([O.fr_from_int(0 if sparse and i % 7 == 0 else i * 31 + 7) for i in range(n * m)])
    dv = H.DeviceVec.upload(sc)
    H.test_delayed_scale(dv, O.fr_from_int(2), 100)
    start = 17 if ragged else 0
    if ragged:
        slices = []
        for b in range(m):
            part = sc[b * n + start:(b + 1) * n]
            slices.append((b * n + start, start, n - start, int(np.count_nonzero(part.any(axis=1)))))
        got = ctx.device_ragged(dv, slices)
    else:
        got = ctx.device_batch(dv, 0, n, m)
    two = O.fr_from_int(2)
    for b in range(m):
        doubled = np.stack([O.fr_op("mul", x, two) for x in sc[b * n + start:(b + 1) * n]])
        assert O.g1_eq(got[b], closed_form(O, ts[start:], doubled)), (ragged, sparse, b)
    ctx.free()
    dv.free()


def test_scrambled_msm_waits_for_delayed_mask_producer(H, O):
    """lib.rs:5088-5108: the masked scalars start zeroed and their producer
    is held back 100 ms; the scrambled MSM must return MSM(P, s + mask)."""
    n, m = 4096, 3
    ts, pts = _reference_points(O, n)
    ctx = H.MsmContext(pts)
    sc = O.random_fr(n * m, 61)
    mask = O.random_fr(n, 62)
    dv = H.DeviceVec.upload(sc)
    mdv = H.DeviceVec.upload(mask)
    try:
        H.test_mask_delay(100)
        got = ctx.device_scrambled_batch(dv, 0, n, m, mdv)
    finally:
        H.test_mask_delay(0)
    for b in range(m):
        s_plus = np.stack([O.fr_op("add", sc[b * n + j], mask[j]) for j in range(n)])
        assert O.g1_eq(got[b], closed_form(O, ts, s_plus)), b
    ctx.free()
    dv.free()
    mdv.free()


def test_msm_context_window_bits_and_autotune(H, O):
    """A context made without a config is autotuned at creation (window of
    the size rule or one bit either side, chunks of 64 or 32); a caller's
    window_bits is used as given (tables capped at 17, plain windows as
    given, .cu:50-110 validity); every choice gives the closed-form MSM."""
    n = 1 << 16
    ts = O.random_fr(n, 71)
    pts = points_for(O, ts)
    sc = O.random_fr(2 * n, 72)
    auto = H.MsmContext(pts)
    w, L, ms = auto.tuning()
    assert w in (11, 12, 13) and L in (32, 64) and ms > 0
    want = [closed_form(O, ts, sc[i * n:(i + 1) * n]) for i in range(2)]
    got = auto.batch(sc, n, 2)
    assert all(O.g1_eq(a, b) for a, b in zip(got, want))
    auto.free()
    for wb in (10, 13, 17, 18):
        cfg = H.default_config(n)
        cfg.window_bits = wb
        ctx = H.MsmContext(pts, config=cfg)
        w, L, ms = ctx.tuning()
        assert w == min(wb, 17) and ms == 0
        got = ctx.batch(sc, n, 2)
        assert all(O.g1_eq(a, b) for a, b in zip(got, want)), wb
        # a short MSM runs plain windows of the configured width
        short = ctx.batch(sc[:300], 300, 1)[0]
        assert O.g1_eq(short, closed_form(O, ts[:300], sc[:300])), wb
        ctx.free()
    cfg = H.default_config(n)
    cfg.window_bits = 9                       # outside the reference's valid range
    with pytest.raises(H.MidnightHipError):
        H.MsmContext(pts, config=cfg)
