"""GPU parity: batch plans (ragged iNTT + batched MSM, graph replay,
autotuned MSM configuration) through the C ABI.

Every replay of a captured plan is checked against the oracle: the
coefficients after r runs equal r oracle inverse NTTs of the input, and
each instance's MSM equals the closed form [sum s_i t_i]G over the
context's points [t_i]G.  Between replays the test runs its own eager
device work (another device vector's transforms and MSMs) and
synchronises, so that a graph that depended on state left by eager work
would show it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _layout(ks):
    offs, total = [], 0
    for k in ks:
        offs.append(total)
        total += 1 << k
    return offs, total


def _context(H, O, log_n, seed):
    ts = O.random_fr(1 << log_n, seed)
    dv_t = H.DeviceVec.upload(ts)
    ctx = H.MsmContext.from_multiples(dv_t)
    dv_t.free()
    return ctx, ts


def _check_instances(O, got, coeffs, offs, ks, ts, tag):
    for i, (o, k) in enumerate(zip(offs, ks)):
        n = 1 << k
        want = O.g1_mul(O.generator(), O.inner_product(coeffs[o:o + n], ts[:n]))
        assert O.g1_eq(got[i], want), "%s: instance %d (k=%d)" % (tag, i, k)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_batch_plan_replays(H, O, flags):
    ks = [10, 12, 11, 10, 13, 12, 10, 14, 11, 11]
    offs, total = _layout(ks)
    ctx, ts = _context(H, O, 14, 900 + flags)
    data = O.random_fr(total, 901 + flags)
    dv = H.DeviceVec.upload(data)
    other = H.DeviceVec.upload(O.random_fr(total, 902 + flags))
    plan = H.BatchPlan(ctx, dv, offs, ks, flags)
    info = plan.info()
    assert info["graph"] == bool(flags & 1)
    cur = data.copy()
    for rep in range(3):
        got = plan.run()
        for o, k in zip(offs, ks):
            cur[o:o + (1 << k)] = O.ifft(cur[o:o + (1 << k)], k)
        assert np.array_equal(dv.download(), cur), "replay %d: coefficients" % rep
        _check_instances(O, got, cur, offs, ks, ts, "flags %d replay %d" % (flags, rep))
        # eager work of the program's own between replays
        other.ntt_multi(offs, ks, True)
        ctx.device_multi(other, offs, [1 << k for k in ks])
    plan.free()
    other.free()
    dv.free()
    ctx.free()


def test_batch_plan_plain_windows_and_errors(H, O):
    # a context below 2^10 points has no precomputed table: plain windows
    ks = [8, 9, 7, 9]
    offs, total = _layout(ks)
    ctx, ts = _context(H, O, 9, 910)
    data = O.random_fr(total, 911)
    dv = H.DeviceVec.upload(data)
    plan = H.BatchPlan(ctx, dv, offs, ks)
    assert not plan.info()["precomputed"]
    got = plan.run()
    cur = data.copy()
    for o, k in zip(offs, ks):
        cur[o:o + (1 << k)] = O.ifft(cur[o:o + (1 << k)], k)
    _check_instances(O, got, cur, offs, ks, ts, "plain")
    plan.free()
    with pytest.raises(H.MidnightHipError):
        H.BatchPlan(ctx, dv, [0, 100], [8, 8])          # overlapping instances
    with pytest.raises(H.MidnightHipError):
        H.BatchPlan(ctx, dv, [0], [10])                 # longer than the context
    with pytest.raises(H.MidnightHipError):
        H.BatchPlan(ctx, dv, [total - 128], [8])        # past the device vector
    with pytest.raises(H.MidnightHipError):
        dv.ntt_multi([0, 128], [8, 8], True)            # overlapping transforms
    dv.free()
    ctx.free()


def test_batch_plan_benchmark_shape(H, O):
    """The batched workload's shape (k = floor(10 + 9u) per instance seed,
    laid out by size) at 48 instances over a 2^16 context: every MSM equals
    the eager batched entry point's, and the first instances the closed
    form."""
    count = 48
    ks = sorted(min(16, max(10, int(10 + 7 * np.random.default_rng(5000 + j).random()))) for j in range(count))
    offs, total = _layout(ks)
    ctx, ts = _context(H, O, 16, 920)
    data = O.random_fr(total, 921)
    dv = H.DeviceVec.upload(data)
    ref = H.DeviceVec.upload(data)
    plan = H.BatchPlan(ctx, dv, offs, ks)
    for rep in range(2):
        got = plan.run()
        ref.ntt_multi(offs, ks, True)
        want = ctx.device_multi(ref, offs, [1 << k for k in ks])
        for i in range(count):
            assert O.g1_eq(got[i], want[i]), "replay %d instance %d" % (rep, i)
    assert np.array_equal(dv.download(), ref.download())
    cur = ref.download()
    _check_instances(O, got[:4], cur, offs[:4], ks[:4], ts, "closed form")
    plan.free()
    ref.free()
    dv.free()
    ctx.free()
