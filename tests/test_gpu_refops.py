"""GPU parity of the reference's device-vector, polynomial, multi-open,
inversion, lookup-product and trash entry points (bls12_381_g1_msm.cu:
2878-3179, 3234, 3558-4110, 4451-4634, 5139, 5667, 5849, 6582, 6617) through
the C ABI.

Bar: limb-identical to the oracle (oracle/poly.c eval / kate division,
oracle/quotient.c lookup products and graph rows, Fr inversion) or, for
pure data movement, to numpy; errors where the reference rejects its
arguments.
"""
import numpy as np
import pytest

import numerator_graph as G

pytestmark = pytest.mark.gpu


def _lincomb(O, polys, scalars, n):
    """sum_j scalars[j] * polys[j] coefficient by coefficient (oracle inner products)"""
    m = len(scalars)
    p = np.ascontiguousarray(polys).reshape(m, n, 4)
    return np.array([O.inner_product(np.ascontiguousarray(p[:, i]), scalars) for i in range(n)])


def test_device_vec_scatter_concat_gather(H, O):
    v = O.random_fr(1000, 800)
    dv = H.DeviceVec.upload(v)
    idx = [3, 999, 0, 500, 3]
    vals = O.random_fr(len(idx), 801)
    dv.scatter_host(idx, vals)
    want = v.copy()
    for i, x in zip(idx, vals):
        want[i] = x
    assert np.array_equal(dv.download(), want)
    with pytest.raises(H.MidnightHipError):
        dv.scatter_host([1000], vals[:1])
    dv.scatter_host([], [])                              # nothing to do
    sfx = O.random_fr(37, 802)
    cat = dv.concat_host(sfx)
    assert np.array_equal(cat.download(), np.concatenate([want, sfx]))
    cat0 = dv.concat_host([])
    assert np.array_equal(cat0.download(), want)
    other = H.DeviceVec.upload(O.random_fr(64, 803))
    many = H.DeviceVec.concat_many_host([dv, other, dv], sfx)
    assert np.array_equal(many.download(), np.concatenate([want, other.download(), want, sfx]))
    polys = O.random_fr(8 * 100, 804)
    pv = H.DeviceVec.upload(polys)
    g = pv.gather_polys(100, [7, 0, 7, 3])
    assert np.array_equal(g.download(), np.concatenate([polys[700:800], polys[:100], polys[700:800],
                                                        polys[300:400]]))
    with pytest.raises(H.MidnightHipError):
        pv.gather_polys(100, [8])
    with pytest.raises(H.MidnightHipError):
        pv.gather_polys(300, [0])                        # 800 is not a multiple of 300
    for x in (dv, cat, cat0, other, many, pv, g):
        x.free()


def test_device_vec_nonzero_suffixes(H, O):
    n, count = 300, 4
    polys = O.random_fr(n * count + 10, 810)
    polys[10:10 + 17] = 0                                # poly 0: first 17 zero
    polys[10 + n:10 + 2 * n] = 0                         # poly 1: all zero
    polys[10 + 2 * n + 5] = 0                            # poly 2: an inner zero
    dv = H.DeviceVec.upload(polys)
    got, starts, counts = dv.download_nonzero_suffixes(10, n, count)
    assert starts == [17, n, 0, 0]
    assert counts == [n - 17, 0, n - 1, n]
    want = np.concatenate([polys[10 + j * n + s:10 + (j + 1) * n] for j, s in enumerate(starts)])
    assert np.array_equal(got, want)
    with pytest.raises(H.MidnightHipError):
        dv.download_nonzero_suffixes(11, n, count)       # past the end
    dv.free()


def test_poly_lincomb_and_eval(H, O):
    n, m = 700, 5
    polys = O.random_fr(n * m, 820)
    sc = O.random_fr(m, 821)
    want = _lincomb(O, polys, sc, n)
    assert np.array_equal(H.fr_poly_lincomb(polys, n, sc), want)
    dv = H.DeviceVec.upload(polys)
    lc = dv.poly_lincomb(n, m, sc)
    assert np.array_equal(lc.download(), want)
    x = O.random_fr(1, 822)[0]
    ev = dv.eval_polys(n, m, x)
    for j in range(m):
        assert np.array_equal(ev[j], O.poly_eval(polys[j * n:(j + 1) * n], x)), j
    with pytest.raises(H.MidnightHipError):
        dv.poly_lincomb(n, m, sc[:4])                    # scalar count mismatch
    with pytest.raises(H.MidnightHipError):
        dv.eval_polys(n + 1, m, x)                       # shape mismatch
    lc.free()
    dv.free()


@pytest.mark.parametrize("m", [16, 17, 40])
def test_poly_lincomb_argument_and_table_paths(H, O, m):
    """Up to 16 polynomials travel in the kernel arguments, more through an
    uploaded pointer / scalar table: both sides of the boundary."""
    n = 257
    polys = O.random_fr(n * m, 850 + m)
    sc = O.random_fr(m, 900 + m)
    want = _lincomb(O, polys, sc, n)
    assert np.array_equal(H.fr_poly_lincomb(polys, n, sc), want)
    dv = H.DeviceVec.upload(polys)
    lc = dv.poly_lincomb(n, m, sc)
    assert np.array_equal(lc.download(), want)
    lc.free()
    dv.free()


def test_device_vectors_cross_lanes(H, O):
    """A device vector returned on one lane is complete: consumed at once on
    another lane (the reference synchronises before releasing a handle,
    bls12_381_g1_msm.cu:3269) it holds the same values as on its own lane."""
    n, m = 1 << 17, 12
    polys = O.random_fr(n * m, 870)
    sc = O.random_fr(m, 871)
    try:
        H.set_lane(0)
        dv0 = H.DeviceVec.upload(polys)
        want = dv0.poly_lincomb(n, m, sc).download()
        rows = [0, 1, n // 2, n - 1]
        p = polys.reshape(m, n, 4)
        for i in rows:
            assert np.array_equal(want[i], O.inner_product(np.ascontiguousarray(p[:, i]), sc)), i
        zeta = O.fr_from_int(1)
        want_ext = H.device_coeff_to_extended_from_device_batch(dv0, n, m, 19, zeta, zeta).download()
        q_dv0 = H.DeviceVec.upload(polys[:3 * n])
        pts = O.random_fr(3, 872)
        want_f = H.fr_multiopen_f_poly_device(q_dv0, n, [0, 1, 2, 3], pts, sc[:3]).download()
        for rep in range(2):
            H.set_lane(1)
            dv1 = H.DeviceVec.upload(polys)
            lc = dv1.poly_lincomb(n, m, sc)
            ext = H.device_coeff_to_extended_from_device_batch(dv1, n, m, 19, zeta, zeta)
            q_dv1 = H.DeviceVec.upload(polys[:3 * n])
            f = H.fr_multiopen_f_poly_device(q_dv1, n, [0, 1, 2, 3], pts, sc[:3])
            H.set_lane(0)
            assert np.array_equal(lc.download(), want), rep
            assert np.array_equal(ext.download(), want_ext), rep
            assert np.array_equal(f.download(), want_f), rep
            for x in (lc, ext, f, q_dv1, dv1):
                x.free()
        dv0.free()
        q_dv0.free()
    finally:
        H.set_lane(0)


def _multiopen_case(O, n, seed):
    polys = O.random_fr(n * 6, seed)
    q_offsets = [0, 3, 4, 7]
    q_poly_indices = [0, 2, 5, 1, 3, 4, 0]
    q_coefficients = O.random_fr(len(q_poly_indices), seed + 1)
    points = O.random_fr(3, seed + 2)
    query_poly = [0, 1, 2, 5, 3, 0]
    query_pt = [0, 0, 1, 2, 1, 2]
    return polys, q_offsets, q_poly_indices, q_coefficients, points, query_poly, query_pt


def _want_multiopen(O, n, polys, q_offsets, q_poly_indices, q_coefficients, points, query_poly, query_pt):
    qs = []
    for k in range(len(q_offsets) - 1):
        items = range(q_offsets[k], q_offsets[k + 1])
        sel = np.concatenate([polys[q_poly_indices[t] * n:(q_poly_indices[t] + 1) * n] for t in items])
        qs.append(_lincomb(O, sel, q_coefficients[q_offsets[k]:q_offsets[k + 1]], n))
    evals = [O.poly_eval(polys[p * n:(p + 1) * n], points[x]) for p, x in zip(query_poly, query_pt)]
    return np.concatenate(qs), np.array(evals)


def test_multiopen_prepare_host_and_resident(H, O):
    n = 1 << 10
    case = _multiopen_case(O, n, 830)
    want_q, want_ev = _want_multiopen(O, n, *case)
    q, ev = H.fr_multiopen_prepare(case[0], n, *case[1:])
    assert np.array_equal(q, want_q)
    assert np.array_equal(ev, want_ev)
    # resident: the same six polynomials spread over three device vectors
    polys = case[0]
    dyn = H.DeviceVec.upload(np.concatenate([polys[2 * n:3 * n], polys[0:n]]))   # 0 -> poly 2, 1 -> poly 0
    fix = H.DeviceVec.upload(np.concatenate([polys[n:2 * n], polys[3 * n:4 * n], polys[5 * n:6 * n]]))
    ext = H.DeviceVec.upload(polys[4 * n:5 * n])
    refs = [(0, 1), (1, 0), (0, 0), (1, 1), (2, 0), (1, 2)]   # logical poly j
    q_host, q_dv, ev2 = H.fr_multiopen_prepare_device(dyn, fix, ext, n, refs, *case[1:])
    assert np.array_equal(q_host, want_q)
    assert np.array_equal(q_dv.download(), want_q)
    assert np.array_equal(ev2, want_ev)
    with pytest.raises(H.MidnightHipError):
        H.fr_multiopen_prepare_device(dyn, fix, ext, n, refs[:5] + [(1, 3)], *case[1:])   # index out of range
    with pytest.raises(H.MidnightHipError):
        H.fr_multiopen_prepare(case[0], n, [0, 3, 2, 7], *case[2:])                        # non-monotonic
    for x in (dyn, fix, ext, q_dv):
        x.free()


def test_multiopen_f_and_pi(H, O):
    n = 1 << 11
    q_count = 3
    qs = O.random_fr(n * q_count, 840)
    points = O.random_fr(4, 841)
    points[3] = 0                                        # a point at zero divides by X
    point_offsets = [0, 1, 3, 4]
    scalars = O.random_fr(q_count, 842)
    q_dv = H.DeviceVec.upload(qs)
    f = H.fr_multiopen_f_poly_device(q_dv, n, point_offsets, points, scalars)
    divided = []
    for k in range(q_count):
        d = qs[k * n:(k + 1) * n]
        for p in range(point_offsets[k], point_offsets[k + 1]):
            d = O.kate_division(d, points[p])
        divided.append(d)
    want_f = _lincomb(O, np.concatenate(divided), scalars, n)
    assert np.array_equal(f.download(), want_f)
    s2 = O.random_fr(q_count + 1, 843)
    x3 = O.random_fr(1, 844)[0]
    pi = H.fr_multiopen_pi_device(q_dv, f, n, s2, x3)
    final = _lincomb(O, np.concatenate([qs, want_f]), s2, n)
    want_pi = O.kate_division(final, x3)[:n - 1]
    got_pi = pi.download()
    assert got_pi.shape[0] == n - 1
    assert np.array_equal(got_pi, want_pi)
    with pytest.raises(H.MidnightHipError):
        H.fr_multiopen_f_poly_device(q_dv, n, [0, 1, 3], points[:3], scalars)   # q count != scalars
    for x in (q_dv, f, pi):
        x.free()


def test_invert_assign(H, O):
    v = O.random_fr(5000, 850)
    v[[0, 17, 4999]] = 0
    got = H.fr_invert_assign(v)
    for i in (0, 1, 2, 17, 31, 32, 33, 2500, 4998, 4999):
        assert np.array_equal(got[i], O.fr_op("inv", v[i])), i
    # x * x^-1 = 1 everywhere it is defined
    one = O.fr_from_int(1)
    for i in range(0, 5000, 97):
        if v[i].any():
            assert np.array_equal(O.fr_op("mul", v[i], got[i]), one), i


@pytest.mark.parametrize("n,count,bf", [(64, 1, 5), (4096, 3, 5), (1 << 14, 2, 0)])
def test_lookup_products(H, O, n, count, bf):
    ci, ct, pi, pt = (O.random_fr(n * count, 860 + k) for k in range(4))
    beta, gamma = O.random_fr(2, 865)
    pi[7] = O.fr_op("neg", beta)                          # zero denominator on row 7 of lookup 0
    blinds = O.random_fr(max(count * bf, 1), 866)[:count * bf]
    want = O.lookup_products(ci, ct, pi, pt, n, count, beta, gamma, blinds, bf)
    dv = H.fr_lookup_products(ci, ct, pi, pt, n, count, beta=beta, gamma=gamma, blinds=blinds,
                              blinding_factors=bf)
    assert np.array_equal(dv.download(), want)
    trace = H.DeviceVec.upload(np.concatenate([ci, ct, pi, pt]))
    dv2 = H.fr_lookup_products_device(trace, n, count, beta=beta, gamma=gamma, blinds=blinds, blinding_factors=bf)
    assert np.array_equal(dv2.download(), want)
    with pytest.raises(H.MidnightHipError):
        H.fr_lookup_products_device(trace, n + 1, count, beta=beta, gamma=gamma, blinds=blinds,
                                    blinding_factors=bf)
    for x in (dv, dv2, trace):
        x.free()


def test_trace_context_trash(H, O):
    n_ext, n = 4096, 1024
    fixed_lag = O.random_fr(3 * n, 870)
    advice = O.random_fr(5 * n, 871)
    challenges = O.random_fr(2, 872)
    trash_ch = O.random_fr(1, 873)[0]
    trash = [G.RawGraphBuffers(G.trash_graph([("advice", 2, 0), ("prod", ("advice", 3, 1), ("challenge", 0))])),
             G.RawGraphBuffers(G.trash_graph([("fixed", 1, -1), ("advice", 0, 0)]))]
    filler = O.random_fr(n_ext * 3, 874)
    ctx = H.NumeratorContext(
        n_ext, n, fixed=filler, fixed_lagrange=fixed_lag, l0=filler[:n_ext], l_last=filler[:n_ext],
        l_active=filler[:n_ext], permutation_columns=[], permutation_pk_cosets=None, permutation_pk_lagrange=None,
        custom_graph=G.RawGraphBuffers(G.custom_gates_graph([("advice", 0, 0)])), omega=O.omega(12),
        trace_omega=O.omega(10), g_coset=O.fr_from_int(7), delta_const=O.fr_from_int(1), rot_scale=4,
        last_rotation=-6, chunk_len=3, trash_graphs=trash, trash_selector_fixed_indices=[0, 1])
    adv = H.DeviceVec.upload(advice)
    out = ctx.trace_trash(adv, 5, None, 0, challenges, trash_ch)
    got = out.download()
    for t, g in enumerate(trash):
        want = O.graph_rows(n, g, rot_scale=1, fixed=fixed_lag, advice=advice, challenges=challenges,
                            trash_challenge=trash_ch)
        assert np.array_equal(got[t * n:(t + 1) * n], want), t
    with pytest.raises(H.MidnightHipError):
        ctx.trace_trash(adv, 4, None, 0, challenges, trash_ch)   # advice length mismatch
    out.free()
    adv.free()
    ctx.free()


def test_sort_canonical(H, O):
    v = O.random_fr(20000, 880)
    v[100:200] = v[:100]                                  # duplicates
    v[5] = 0
    assert np.array_equal(H.fr_sort_canonical(v), O.sort_canonical(v))
    small = O.random_fr(1, 881)
    assert np.array_equal(H.fr_sort_canonical(small), small)


def _lookup_case(O, n, count, seed, usable, table_rows=300):
    """inputs drawn (with heavy repetition) from the usable rows of each table"""
    tables = O.random_fr(n * count, seed)
    rng = np.random.default_rng(seed)
    rows = min(table_rows, usable)
    inputs = np.concatenate([tables[l * n + rng.integers(0, rows, n)] for l in range(count)])
    return inputs, tables


def _want_trace(O, inputs, tables, n, count, usable, ib, tb):
    bc = n - usable
    pin, ptab = [], []
    for l in range(count):
        a, t = O.lookup_permute(inputs[l * n:l * n + usable], tables[l * n:l * n + usable], usable)
        pin.append(np.concatenate([a, ib[l * bc:(l + 1) * bc]]))
        ptab.append(np.concatenate([t, tb[l * bc:(l + 1) * bc]]))
    return np.concatenate([inputs, tables] + pin + ptab)


@pytest.mark.parametrize("n,count", [(256, 1), (1 << 14, 3)])
def test_lookup_permute_batch(H, O, n, count):
    usable = n - 6
    inputs, tables = _lookup_case(O, n, count, 890 + count, usable)
    ib = O.random_fr(count * 6, 893)
    tb = O.random_fr(count * 6, 894)
    dv = H.fr_lookup_permute_batch(inputs, tables, n, count, usable, ib, tb)
    assert np.array_equal(dv.download(), _want_trace(O, inputs, tables, n, count, usable, ib, tb))
    dv.free()
    bad = inputs.copy()
    bad[10] = O.random_fr(1, 895)[0]                        # not in the table
    with pytest.raises(H.MidnightHipError) as e:
        H.fr_lookup_permute_batch(bad, tables, n, count, usable, ib, tb)
    assert e.value.code == -1


def test_trace_context_lookup_permute(H, O):
    n_ext, n = 4096, 1024
    usable = n - 6
    theta = O.random_fr(1, 900)[0]
    fixed_lag = O.random_fr(3 * n, 901)
    advice = O.random_fr(5 * n, 902)
    # lookup 0: (advice 0, advice 1) in (fixed 0, fixed 1); lookup 1: advice 2 in fixed 2
    rng = np.random.default_rng(903)
    rows = rng.integers(0, 200, n)
    advice[0:n] = fixed_lag[rows]
    advice[n:2 * n] = fixed_lag[n + rows]
    advice[2 * n:3 * n] = fixed_lag[2 * n + rng.integers(0, 50, n)]
    inputs = [G.RawGraphBuffers(G.compression_graph([("advice", 0, 0), ("advice", 1, 0)])),
              G.RawGraphBuffers(G.compression_graph([("advice", 2, 0)]))]
    tables = [G.RawGraphBuffers(G.compression_graph([("fixed", 0, 0), ("fixed", 1, 0)])),
              G.RawGraphBuffers(G.compression_graph([("fixed", 2, 0)]))]
    filler = O.random_fr(n_ext * 3, 904)
    ctx = H.NumeratorContext(
        n_ext, n, fixed=filler, fixed_lagrange=fixed_lag, l0=filler[:n_ext], l_last=filler[:n_ext],
        l_active=filler[:n_ext], permutation_columns=[], permutation_pk_cosets=None, permutation_pk_lagrange=None,
        custom_graph=G.RawGraphBuffers(G.custom_gates_graph([("advice", 0, 0)])), omega=O.omega(12),
        trace_omega=O.omega(10), g_coset=O.fr_from_int(7), delta_const=O.fr_from_int(1), rot_scale=4,
        last_rotation=-6, chunk_len=3, lookup_graphs=inputs, lookup_input_graphs=inputs,
        lookup_table_graphs=tables)
    adv = H.DeviceVec.upload(advice)
    ib = O.random_fr(2 * 6, 905)
    tb = O.random_fr(2 * 6, 906)
    out = ctx.trace_lookup_permute(adv, 5, None, 0, None, theta, usable, ib, tb)
    ci = np.concatenate([O.graph_rows(n, g, theta=theta, fixed=fixed_lag, advice=advice) for g in inputs])
    ct = np.concatenate([O.graph_rows(n, g, theta=theta, fixed=fixed_lag, advice=advice) for g in tables])
    assert np.array_equal(out.download(), _want_trace(O, ci, ct, n, 2, usable, ib, tb))
    # the products of the permuted trace close to one at the last usable row
    beta, gamma = O.random_fr(2, 907)
    z = H.fr_lookup_products_device(out, n, 2, beta=beta, gamma=gamma, blinds=O.random_fr(10, 908),
                                    blinding_factors=5)
    zz = z.download()
    one = O.fr_from_int(1)
    assert np.array_equal(zz[usable], one) and np.array_equal(zz[n + usable], one)
    for x in (z, out, adv):
        x.free()
    ctx.free()
