"""Pinning of the numerator oracle (oracle/quotient.c) and of the host-side
graph compiler (numerator_graph.py), without a GPU.

- The compiled graph of the benchmark gate is the one the reference's
  builder produces (hand-derived from evaluation.rs:3685-3786 and the slot
  planner evaluation.rs:419-545).
- Graph evaluation equals a direct evaluation of the expression trees the
  graphs were compiled from (a separate tree walker here).
- On a satisfying synthetic circuit the whole numerator (gate, permutation
  over three sets) is zero on every row of the base domain; a changed
  witness cell makes exactly the rows it touches nonzero.
- On the extended coset, numerator / (X^n - 1) interpolates to a
  polynomial of degree < (n - 1)(d - 1): its top d - 1 coefficients vanish.
"""
import random

import numpy as np
import pytest

import numerator_graph as G
import oracle as O
import synth_circuit as S

R = G.FR_MODULUS


def test_gate_graph_matches_reference_builder():
    g = G.custom_gates_graph([S.gate_expression()])
    assert g.constants == [0, 1, 2]
    assert g.rotations == [0]
    raw = G.RawGraphBuffers(g)
    assert raw.num_intermediates == 3
    I = lambda s: (G.INTER, s, 0)
    expect = [
        (G.OP_STORE, 0, 0, 0, (G.FIXED, 0, 0), (0, 0, 0)),
        (G.OP_STORE, 1, 0, 0, (G.ADVICE, 0, 0), (0, 0, 0)),
        (G.OP_STORE, 2, 0, 0, (G.ADVICE, 1, 0), (0, 0, 0)),
        (G.OP_MUL, 2, 0, 0, I(1), I(2)),
        (G.OP_STORE, 1, 0, 0, (G.ADVICE, 2, 0), (0, 0, 0)),
        (G.OP_ADD, 1, 0, 0, I(2), I(1)),
        (G.OP_STORE, 2, 0, 0, (G.ADVICE, 3, 0), (0, 0, 0)),
        (G.OP_SUB, 2, 0, 0, I(1), I(2)),
        (G.OP_MUL, 2, 0, 0, I(0), I(2)),
        (G.OP_HORNER, 2, 0, 1, (G.PREV, 0, 0), (G.Y, 0, 0)),
    ]
    assert raw.ops == expect
    assert raw.parts == [I(2)]


def test_builder_simplifications():
    g = G.GraphEvaluator()
    a = ("advice", 0, 0)
    assert g.add_expression(("prod", ("constant", 0), a)) == G.ZERO_CONST
    s = g.add_expression(("prod", ("constant", 1), a))
    assert g.calculations[-1][0] == ("store", (G.ADVICE, 0, 0))
    assert g.add_expression(("prod", a, ("constant", 2))) == g.add_calculation(("double", s))
    assert g.add_expression(("prod", a, a)) == g.add_calculation(("square", s))
    assert g.add_expression(("neg", ("constant", 5))) == (G.CONST, 3, 0) and g.constants[3] == R - 5
    assert g.add_expression(("sum", ("constant", 0), ("neg", a))) == g.add_calculation(("negate", s))
    # a repeated sub-expression is computed once
    n0 = len(g.calculations)
    g.add_expression(("prod", a, ("advice", 1, 1)))
    g.add_expression(("prod", ("advice", 1, 1), a))
    assert len(g.calculations) == n0 + 2


def _walk(e, row, n, rot_scale, cols, challenges):
    """Direct evaluation of an expression tree at one row."""
    t = e[0]
    if t == "constant":
        return e[1] % R
    if t in ("fixed", "advice", "instance"):
        c = cols[t][e[1]]
        return c[(row + e[2] * rot_scale) % n]
    if t == "challenge":
        return challenges[e[1]]
    if t == "neg":
        return (-_walk(e[1], row, n, rot_scale, cols, challenges)) % R
    if t == "sum":
        return (_walk(e[1], row, n, rot_scale, cols, challenges) + _walk(e[2], row, n, rot_scale, cols, challenges)) % R
    if t == "prod":
        return _walk(e[1], row, n, rot_scale, cols, challenges) * _walk(e[2], row, n, rot_scale, cols, challenges) % R
    if t == "scaled":
        return _walk(e[1], row, n, rot_scale, cols, challenges) * e[2] % R
    raise ValueError(t)


def _expressions():
    a, b, c, d = (("advice", i, 0) for i in range(4))
    a_next, a_prev3 = ("advice", 0, 1), ("advice", 0, -3)
    q, s1 = ("fixed", 0, 0), ("fixed", 1, -1)
    inst = ("instance", 0, 2)
    return [
        S.gate_expression(),
        ("sum", ("prod", a_next, a_next), ("neg", ("prod", ("constant", 2), b))),
        ("scaled", ("sum", a_prev3, ("neg", ("constant", 7))), 123456789),
        ("prod", s1, ("sum", inst, ("challenge", 1))),
        ("sum", ("neg", ("sum", c, d)), ("prod", ("constant", 1), ("prod", q, ("constant", 0)))),
        ("prod", ("sum", a, b), ("sum", b, a)),
    ]


@pytest.mark.parametrize("rot_scale", [1, 2])
def test_graph_eval_equals_direct_evaluation(rot_scale):
    n = 64
    rnd = random.Random(7)
    cols = {"advice": [[rnd.randrange(R) for _ in range(n)] for _ in range(4)],
            "fixed": [[rnd.randrange(R) for _ in range(n)] for _ in range(2)],
            "instance": [[rnd.randrange(R) for _ in range(n)]]}
    ch = [rnd.randrange(R) for _ in range(2)]
    y = rnd.randrange(R)
    prev = [rnd.randrange(R) for _ in range(n)]
    exprs = _expressions()
    graph = G.RawGraphBuffers(G.custom_gates_graph(exprs))
    flat = lambda cs: S.mont([x for c in cs for x in c])
    got = O.graph_rows(n, graph, y=S.mont([y])[0], rot_scale=rot_scale, fixed=flat(cols["fixed"]),
                       advice=flat(cols["advice"]), instance=flat(cols["instance"]),
                       challenges=S.mont(ch), previous=S.mont(prev))
    got = S.ints(got)
    for row in range(n):
        v = prev[row]
        for e in exprs:
            v = (v * y + _walk(e, row, n, rot_scale, cols, ch)) % R
        assert got[row] == v, row


def test_lookup_and_trash_graph_shapes():
    inputs = [("advice", 0, 0), ("prod", ("advice", 1, 0), ("fixed", 0, 0))]
    tables = [("fixed", 1, 0), ("fixed", 2, 0)]
    g = G.lookup_graph(inputs, tables)
    n = 16
    rnd = random.Random(3)
    adv = [[rnd.randrange(R) for _ in range(n)] for _ in range(2)]
    fix = [[rnd.randrange(R) for _ in range(n)] for _ in range(3)]
    beta, gamma, theta = (rnd.randrange(R) for _ in range(3))
    got = S.ints(O.graph_rows(n, G.RawGraphBuffers(g), beta=S.mont([beta])[0], gamma=S.mont([gamma])[0],
                              theta=S.mont([theta])[0], advice=S.mont([x for c in adv for x in c]),
                              fixed=S.mont([x for c in fix for x in c])))
    for i in range(n):
        ci = (adv[0][i] * theta + adv[1][i] * fix[0][i]) % R
        ct = (fix[1][i] * theta + fix[2][i]) % R
        assert got[i] == (ci + beta) * (ct + gamma) % R
    tg = G.trash_graph([("advice", 0, 0), ("advice", 1, 0), ("constant", 5)])
    tr = rnd.randrange(R)
    got = S.ints(O.graph_rows(n, G.RawGraphBuffers(tg), trash_challenge=S.mont([tr])[0],
                              advice=S.mont([x for c in adv for x in c])))
    for i in range(n):
        assert got[i] == ((adv[0][i] * tr + adv[1][i]) * tr + 5) % R


@pytest.fixture(scope="module")
def circuit():
    return S.Circuit(7, seed=11)


def test_synthetic_circuit_is_satisfied(circuit):
    c = circuit
    assert len(c.cycles) > 8
    for cyc in c.cycles:
        vals = {c.cols[j][i] for j, i in cyc}
        assert len(vals) == 1
    # the last product of the last set closes the argument
    assert c.z[-1][c.u] == 1
    assert c.z[0][0] == 1


def test_numerator_vanishes_on_base_domain(circuit):
    nu = O.numerator(**circuit.base_inputs())
    assert not nu.any()


def test_numerator_flags_broken_gate_row(circuit):
    c = circuit
    inp = c.base_inputs()
    row = next(r for r in range(3, c.u) if all((0, r) not in cyc for cyc in c.cycles))
    adv = inp["advice"].copy()
    adv[0 * c.n + row] = S.mont([c.cols[0][row] + 1])[0]
    inp["advice"] = adv
    nu = O.numerator(**inp)
    assert [i for i in range(c.n) if nu[i].any()] == [row]


def test_numerator_flags_broken_copy(circuit):
    c = circuit
    inp = c.base_inputs()
    j, row = next(cell for cyc in c.cycles for cell in cyc if cell[0] == 4)
    adv = inp["advice"].copy()
    adv[j * c.n + row] = S.mont([c.cols[j][row] + 1])[0]
    inp["advice"] = adv
    nu = O.numerator(**inp)
    assert [i for i in range(c.n) if nu[i].any()] == [row]


def test_numerator_flags_wrong_set_chaining(circuit):
    c = circuit
    inp = c.base_inputs()
    z = inp["permutation_products"].copy().reshape(3, c.n, 4)
    inp["permutation_products"] = np.ascontiguousarray(z[[1, 0, 2]].reshape(-1, 4))
    nu = O.numerator(**inp)
    assert nu[0].any()


def test_numerator_extended_quotient_degree(circuit):
    c = circuit
    inp = c.extended_inputs()
    n, N = c.n, inp["n"]
    assert N == 4 * n
    nu = O.numerator(**inp)
    h = O.mul_repeating(nu, c.t_evals)
    coeffs = O.extended_to_coeff(h, c.domain)
    top = (n - 1) * (S.DEGREE - 1)
    assert not coeffs[top:].any()
    assert coeffs[top - 1].any()
    # a changed advice value breaks divisibility by X^n - 1
    bad = dict(inp)
    adv = inp["advice"].copy()
    adv[5] = O.fr_op("add", adv[5], O.fr_from_int(1))
    bad["advice"] = adv
    coeffs = O.extended_to_coeff(O.mul_repeating(O.numerator(**bad), c.t_evals), c.domain)
    assert coeffs[top:].any()


def test_oracle_permutation_products_match_prover_restatement(circuit):
    """The C grand products equal the circuit builder's (a separate Python
    restatement of permutation/prover.rs), and the last set closes to one
    at the last usable row."""
    c = circuit
    cols = S.mont([x for col in c.cols for x in col])
    sig = S.mont([x for col in c.sigma for x in col])
    z = O.permutation_products(cols, sig, c.n, S.CHUNK, S.mont([c.beta])[0], S.mont([c.gamma])[0],
                               S.mont([c.omega])[0], S.mont(c.delta_pows), S.mont(c.blinds), S.BLINDING)
    assert np.array_equal(z, S.mont([x for zz in c.z for x in zz]))
    assert S.ints(z[2 * c.n + c.u][None, :])[0] == 1


R_MOD = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001


def test_oracle_lookup_products_brute_force(O):
    """orc_lookup_products against Python integers on a tiny input, with a
    zero denominator (a fraction of zero) and the blinded tail."""
    n, count, bf = 8, 2, 2
    vals = [[random.Random(700 + k).randrange(R_MOD) for _ in range(n * count)] for k in range(4)]
    beta, gamma = 11, 13
    vals[2][3] = (R_MOD - beta) % R_MOD        # permuted input + beta == 0 on row 3 of lookup 0
    blinds = [random.Random(710).randrange(R_MOD) for _ in range(count * bf)]
    arrs = [np.array([O.fr_from_int(v) for v in col]) for col in vals]
    got = O.lookup_products(*arrs, n, count, O.fr_from_int(beta), O.fr_from_int(gamma),
                            np.array([O.fr_from_int(b) for b in blinds]), bf)
    for l in range(count):
        z = [1]
        for i in range(n - bf - 1):
            o = l * n + i
            den = (vals[2][o] + beta) * (vals[3][o] + gamma) % R_MOD
            num = (vals[0][o] + beta) * (vals[1][o] + gamma) % R_MOD
            f = num * pow(den, R_MOD - 2, R_MOD) % R_MOD if den else 0
            z.append(z[-1] * f % R_MOD)
        z += blinds[l * bf:(l + 1) * bf]
        assert [O.fr_to_int(got[l * n + i]) for i in range(n)] == z, l
    assert O.fr_to_int(got[4]) == 0   # the product is zero after the zero denominator


def test_oracle_lookup_products_close_on_permutations(O):
    """When the permuted columns are permutations of the compressed ones
    over the usable rows, every lookup's product returns to one at the last
    usable row (the lookup argument's l_last check)."""
    n, count, bf = 64, 3, 5
    u = n - bf - 1
    rng = random.Random(720)
    ci, ct, pi, pt = ([] for _ in range(4))
    for l in range(count):
        a = [rng.randrange(R_MOD) for _ in range(n)]
        s = [rng.randrange(R_MOD) for _ in range(n)]
        pa, ps = a[:u], s[:u]
        rng.shuffle(pa)
        rng.shuffle(ps)
        ci += a
        ct += s
        pi += pa + a[u:]
        pt += ps + s[u:]
    arrs = [np.array([O.fr_from_int(v) for v in col]) for col in (ci, ct, pi, pt)]
    z = O.lookup_products(*arrs, n, count, O.fr_from_int(5), O.fr_from_int(7), O.random_fr(count * bf, 721), bf)
    for l in range(count):
        assert O.fr_to_int(z[l * n]) == 1
        assert O.fr_to_int(z[l * n + u]) == 1, l
        assert O.fr_to_int(z[l * n + u // 2]) != 1


def test_oracle_lookup_permute(O):
    """orc_lookup_permute: a hand-derived case of permute_expression_pair
    (ascending leftovers fill the repeated rows from the last one back), the
    lookup argument's row property on a random case, and the error for an
    input missing from the table."""
    f = O.fr_from_int
    pin, ptab = O.lookup_permute(np.array([f(3), f(1), f(3), f(2), f(3)]),
                                 np.array([f(1), f(2), f(3), f(4), f(5)]), 5)
    assert [O.fr_to_int(x) for x in pin] == [1, 2, 3, 3, 3]
    assert [O.fr_to_int(x) for x in ptab] == [1, 2, 3, 5, 4]
    assert O.lookup_permute(np.array([f(6)] * 5), np.array([f(1), f(2), f(3), f(4), f(5)]), 5) is None
    rng = random.Random(730)
    n = 300
    table = [rng.randrange(R_MOD) for _ in range(n)]
    inputs = [table[rng.randrange(40)] for _ in range(n)]     # heavy repetition, all in the table
    pin, ptab = O.lookup_permute(np.array([f(x) for x in inputs]), np.array([f(x) for x in table]), n)
    pi = [O.fr_to_int(x) for x in pin]
    pt = [O.fr_to_int(x) for x in ptab]
    assert pi == sorted(inputs)
    assert sorted(pt) == sorted(table)
    for r in range(n):
        assert pi[r] == pt[r] or (r > 0 and pi[r] == pi[r - 1]), r
    s = O.sort_canonical(np.array([f(x) for x in table]))
    assert [O.fr_to_int(x) for x in s] == sorted(table)
