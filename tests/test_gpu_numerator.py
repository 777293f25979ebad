"""GPU parity: quotient numerator through the C ABI vs the CPU oracle.

Bar: limb-identical (every value is a fully reduced Montgomery Fr).
Reference entry points: bls12_381_g1_msm.cu:5196 (eval_numerator), :5434
(eval_graph_batch), :5517 / :5662 / :5968 (numerator context create / free /
eval); semantics vendor/midnight-proofs-0.8.1/src/plonk/evaluation.rs:
3114-3528.  The oracle itself is pinned by tests/test_oracle_quotient.py
(base-domain vanishing on a satisfying circuit, quotient degree bound).
"""
import numpy as np
import pytest

import numerator_graph as G
import synth_circuit as S

pytestmark = pytest.mark.gpu


def _delta(O):
    d = O.fr()
    O.lib().orc_fr_delta(O.P(d))
    return d


def gpu_kwargs(O, inp):
    """Oracle numerator inputs -> eval_numerator arguments
    (delta_start = beta * g_coset, delta_const = DELTA)."""
    kw = dict(inp)
    n = kw.pop("n")
    g = kw.pop("g_coset")
    kw["delta_start"] = O.fr_op("mul", kw["beta"], g)
    kw["delta_const"] = _delta(O)
    return n, kw


def run_both(H, O, inp):
    n, kw = gpu_kwargs(O, inp)
    print("numerator n=%d rot_scale=%d" % (n, inp["rot_scale"]), flush=True)
    got = H.eval_numerator(n, **kw)
    want = O.numerator(**inp)
    return got, want


@pytest.mark.parametrize("log_n,seed", [(8, 1), (12, 2), (16, 3)])
def test_numerator_random_inputs(H, O, log_n, seed):
    got, want = run_both(H, O, S.random_inputs(1 << log_n, seed))
    assert np.array_equal(got, want)


def test_numerator_satisfied_circuit(H, O):
    c = S.Circuit(8, seed=5)
    n, kw = gpu_kwargs(O, c.base_inputs())
    got = H.eval_numerator(n, **kw)
    assert not got.any(), "numerator of a satisfying circuit is nonzero on the base domain"
    inp = c.extended_inputs()
    got, want = run_both(H, O, inp)
    assert np.array_equal(got, want)
    coeffs = O.extended_to_coeff(O.mul_repeating(got, c.t_evals), c.domain)
    assert not coeffs[(c.n - 1) * (S.DEGREE - 1):].any()


def test_numerator_ragged_length_and_wrapping(H, O):
    # not a power of two; rotations wrap in both directions
    inp = S.random_inputs(1000, 9, rot_scale=7)
    inp["last_rotation"] = -300
    got, want = run_both(H, O, inp)
    assert np.array_equal(got, want)


def test_numerator_skip_custom_with_initial(H, O):
    n = 4096
    inp = S.random_inputs(n, 12)
    inp["run_custom"] = False
    inp["initial_out"] = O.random_fr(n, 13)
    got, want = run_both(H, O, inp)
    assert np.array_equal(got, want)


def test_numerator_partial_last_set(H, O):
    # 8 columns with chunk_len 3: the last set has two columns; also a set
    # count that exceeds the columns (extra sets contribute z(wX) - z(X))
    for sets in (3, 4):
        inp = S.random_inputs(2048, 20 + sets, sets=sets)
        got, want = run_both(H, O, inp)
        assert np.array_equal(got, want), sets


def _lookup_trash_inputs(O, n, seed):
    inp = S.random_inputs(n, seed)
    lookups = [G.RawGraphBuffers(G.lookup_graph([("advice", 0, 0), ("prod", ("advice", 1, 0), ("fixed", 0, 0))],
                                                [("fixed", 1, 0), ("fixed", 2, -1)])),
               G.RawGraphBuffers(G.lookup_graph([("advice", 4, 1)], [("fixed", 2, 0)]))]
    trash = [G.RawGraphBuffers(G.trash_graph([("advice", 2, 0), ("prod", ("advice", 3, 0), ("challenge", 0))]))]
    extra = O.random_fr(n * 7 + 2, seed + 100)
    inp.update(lookup_product_cosets=extra[:2 * n], lookup_input_cosets=extra[2 * n:4 * n],
               lookup_table_cosets=extra[4 * n:6 * n], lookup_graphs=lookups,
               trash_cosets=extra[6 * n:7 * n], trash_selector_fixed_indices=[1], trash_graphs=trash,
               challenges=extra[7 * n:7 * n + 2])
    return inp


def test_numerator_lookups_and_trash(H, O):
    got, want = run_both(H, O, _lookup_trash_inputs(O, 4096, 31))
    assert np.array_equal(got, want)


def _wide_graph():
    # twelve products kept live until the closing Horner: more slots than
    # the LDS slot file holds, so the global slot file is used
    polys = [("prod", ("advice", i % 5, (i % 3) - 1), ("fixed", i % 3, 0)) for i in range(12)]
    polys += [("sum", ("advice", i % 5, 0), ("constant", i + 3)) for i in range(6)]
    return G.RawGraphBuffers(G.custom_gates_graph(polys))


def test_numerator_wide_graph_global_slots(H, O):
    g = _wide_graph()
    assert g.num_intermediates > 6
    inp = S.random_inputs(5000, 41, graph=g)
    got, want = run_both(H, O, inp)
    assert np.array_equal(got, want)


def test_eval_graph_batch(H, O):
    n = 3000
    inp = _lookup_trash_inputs(O, n, 51)
    graphs = [inp["custom_graph"], _wide_graph()] + inp["lookup_graphs"] + inp["trash_graphs"]
    sc = dict(beta=inp["beta"], gamma=inp["gamma"], theta=inp["theta"],
              trash_challenge=inp["trash_challenge"], y=inp["y"])
    got = H.eval_graph_batch(n, graphs, rot_scale=4, fixed=inp["fixed"], advice=inp["advice"],
                             challenges=inp["challenges"], **sc)
    for gi, g in enumerate(graphs):
        want = O.graph_rows(n, g, rot_scale=4, fixed=inp["fixed"], advice=inp["advice"],
                            challenges=inp["challenges"], **sc)
        assert np.array_equal(got[gi], want), gi


def test_numerator_context_resident(H, O):
    n = 1 << 14
    inp = _lookup_trash_inputs(O, n, 61)
    zeros_lag = O.random_fr(64 * 3 + 64 * 8, 62)
    ctx = H.NumeratorContext(
        n, 64, fixed=inp["fixed"], fixed_lagrange=zeros_lag[:64 * 3], l0=inp["l0"], l_last=inp["l_last"],
        l_active=inp["l_active"], permutation_columns=inp["permutation_columns"],
        permutation_pk_cosets=inp["permutation_pk_cosets"], permutation_pk_lagrange=zeros_lag[64 * 3:],
        custom_graph=inp["custom_graph"], omega=inp["omega"], trace_omega=inp["omega"],
        g_coset=inp["g_coset"], delta_const=_delta(O), rot_scale=inp["rot_scale"],
        last_rotation=inp["last_rotation"], chunk_len=inp["chunk_len"], lookup_graphs=inp["lookup_graphs"],
        lookup_input_graphs=inp["lookup_graphs"], lookup_table_graphs=inp["lookup_graphs"],
        trash_graphs=inp["trash_graphs"], trash_selector_fixed_indices=[1])
    # column vector: advice(5) | products(3) | lookup z(2) | inputs(2) | tables(2) | trash(1)
    cols = np.concatenate([inp["advice"], inp["permutation_products"], inp["lookup_product_cosets"],
                           inp["lookup_input_cosets"], inp["lookup_table_cosets"], inp["trash_cosets"]])
    dv = H.DeviceVec.upload(cols)
    out = ctx.eval(dv, 15, beta=inp["beta"], gamma=inp["gamma"], theta=inp["theta"],
                   trash_challenge=inp["trash_challenge"], y=inp["y"], advice_offset=0, advice_cols=5,
                   challenges=inp["challenges"], permutation_offset=5, permutation_sets=3,
                   lookup_product_offset=8, lookup_input_offset=10, lookup_table_offset=12, lookup_count=2,
                   trash_offset=14, trash_count=1)
    got = out.download()
    want = O.numerator(**inp)
    assert np.array_equal(got, want)
    # a second evaluation with new challenges reuses the context
    inp2 = dict(inp, y=O.random_fr(1, 63)[0])
    out2 = ctx.eval(dv, 15, beta=inp["beta"], gamma=inp["gamma"], theta=inp["theta"],
                    trash_challenge=inp["trash_challenge"], y=inp2["y"], advice_offset=0, advice_cols=5,
                    challenges=inp["challenges"], permutation_offset=5, permutation_sets=3,
                    lookup_product_offset=8, lookup_input_offset=10, lookup_table_offset=12, lookup_count=2,
                    trash_offset=14, trash_count=1)
    assert np.array_equal(out2.download(), O.numerator(**inp2))
    with pytest.raises(H.MidnightHipError):
        ctx.eval(dv, 14, beta=inp["beta"], gamma=inp["gamma"], theta=inp["theta"],
                 trash_challenge=inp["trash_challenge"], y=inp["y"], lookup_count=2, trash_count=1)
    ctx.free()


def test_numerator_argument_errors(H, O):
    inp = S.random_inputs(256, 70)
    n, kw = gpu_kwargs(O, inp)
    kw["run_custom"] = False
    with pytest.raises(H.MidnightHipError, match="initial_out"):
        H.eval_numerator(n, **kw)
    # zero rows: success, nothing written
    assert H.eval_numerator(0, **gpu_kwargs(O, S.random_inputs(256, 71))[1]).shape == (0, 4)
    bad = G.RawGraphBuffers(G.custom_gates_graph([S.gate_expression()]))
    bad.num_intermediates = 1          # targets outside the slot count
    n, kw = gpu_kwargs(O, S.random_inputs(256, 72, graph=bad))
    with pytest.raises(H.MidnightHipError):
        H.eval_numerator(n, **kw)


@pytest.mark.slow
def test_numerator_2p22_extended_rows(H, O):
    """The benchmark circuit's extended domain: 2^22 rows (k = 20, 4x)."""
    n = 1 << 22
    got, want = run_both(H, O, S.random_inputs(n, 81))
    assert np.array_equal(got, want)
