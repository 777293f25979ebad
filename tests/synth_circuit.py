"""Synthetic PLONK circuit for the quotient-numerator tests.

Five advice columns (a, b, c, d, e), three fixed columns (q, s1, s2), one
gate q * (a * b + c - d) = 0, and copy constraints over all eight columns
(random cycles of 2-4 cells), split into permutation sets of
chunk_len = degree - 2 = 3 columns.  Degree 5 (the minimum degree the
benchmark circuit sets), so the extended domain is 4x the base domain;
blinding factors = max(3, 1) + 0 + 1 + 1 = 5 (circuit.rs:2502-2530).

Everything is built here with Python integers, following the reference
prover's steps: permutation keygen sigma values delta^col' * omega^row'
(permutation/keygen.rs), the grand products of permutation/prover.rs:327-515
(denominators beta*sigma + gamma + v, numerators delta^j*omega^i*beta +
gamma + v, shifted prefix product chained over sets, blinded last rows),
and the Lagrange polynomials l0, l_last, l_active of keygen.  The result is
a satisfying witness, on which the numerator must vanish on the base
domain.  Used only by tests.
"""
import random

import numpy as np

import oracle as O
import numerator_graph as G

R = G.FR_MODULUS
DEGREE = 5
CHUNK = DEGREE - 2
BLINDING = 5
ADVICE = 5
FIXED = 3
# permutation columns: advice a..e, then fixed q, s1, s2
PERM_COLUMNS = [(1, i) for i in range(ADVICE)] + [(0, i) for i in range(FIXED)]


def mont(values):
    return G.to_montgomery_limbs(values)


def ints(arr):
    return G.from_montgomery_limbs(arr)


def fr_const(name):
    out = O.fr()
    getattr(O.lib(), "orc_fr_" + name)(O.P(out))
    return ints(out[None, :])[0]


def omega_int(k):
    return ints(O.omega(k)[None, :])[0]


def gate_expression():
    """q * (a * b + c - d), as the expression tree the gate API builds."""
    q = ("fixed", 0, 0)
    a, b, c, d = (("advice", i, 0) for i in range(4))
    return ("prod", q, ("sum", ("sum", ("prod", a, b), c), ("neg", d)))


def gate_graph():
    """the compiled custom-gate graph of the family"""
    return G.RawGraphBuffers(G.custom_gates_graph([gate_expression()]))


class Circuit:
    def __init__(self, k, seed=1, reserved=()):
        rnd = random.Random(seed)
        self.k = k
        n = self.n = 1 << k
        u = self.u = n - BLINDING - 1          # l_last row; rows < u are usable
        self.omega = omega_int(k)
        self.delta = fr_const("delta")
        self.zeta = fr_const("zeta")
        rv = lambda: rnd.randrange(R)

        a = [rv() for _ in range(n)]
        b = [rv() for _ in range(n)]
        c = [rv() for _ in range(n)]
        d = [(a[i] * b[i] + c[i]) % R if i < u else rv() for i in range(n)]
        e = [rv() for _ in range(n)]
        q = [1 if i < u else 0 for i in range(n)]
        s1 = [rv() if i < u else 0 for i in range(n)]
        s2 = [rv() if i < u else 0 for i in range(n)]
        cols = [a, b, c, d, e, q, s1, s2]

        # copy constraints: cycles whose cells hold one value.  Sources are
        # cells of a..d (gate-constrained), targets cells of e, s1, s2 whose
        # values are overwritten; q cells (all 1) form cycles of their own.
        used = set()
        cycles = []
        free_targets = [(col, row) for col in (4, 6, 7) for row in range(u)]
        rnd.shuffle(free_targets)
        for _ in range(max(4, u // 3)):
            src = (rnd.randrange(4), rnd.randrange(u))
            if src in used or src in reserved:
                continue
            size = rnd.randrange(1, 4)
            members = [src]
            while len(members) < size + 1 and free_targets:
                t = free_targets.pop()
                if t in used:
                    continue
                cols[t[0]][t[1]] = cols[src[0]][src[1]]
                members.append(t)
            if len(members) < 2:
                continue
            used.update(members)
            rnd.shuffle(members)
            cycles.append(members)
        qrows = rnd.sample(range(u), min(u, 12))
        for i in range(0, len(qrows) - 1, 3):
            cyc = [(5, r) for r in qrows[i:i + 3]]
            if len(cyc) >= 2:
                cycles.append(cyc)
        self.cols = cols
        self.cycles = cycles
        self.rnd = rnd

        # sigma mapping; cells outside cycles map to themselves
        mapping = {}
        for cyc in cycles:
            for t, cell in enumerate(cyc):
                mapping[cell] = cyc[(t + 1) % len(cyc)]
        ncol = len(PERM_COLUMNS)
        omega_pows = [pow(self.omega, i, R) for i in range(n)]
        delta_pows = [pow(self.delta, j, R) for j in range(ncol)]
        self.sigma = []
        for j in range(ncol):
            col = []
            for i in range(n):
                jj, ii = mapping.get((j, i), (j, i))
                col.append(delta_pows[jj] * omega_pows[ii] % R)
            self.sigma.append(col)
        self.omega_pows = omega_pows
        self.delta_pows = delta_pows

        self.beta, self.gamma, self.theta, self.trash, self.y = (rv() for _ in range(5))
        self.blinds = [rv() for _ in range(BLINDING * self.nsets)]
        self.z = self.products()

        self.l0 = [1 if i == 0 else 0 for i in range(n)]
        self.l_last = [1 if i == u else 0 for i in range(n)]
        self.l_active = [1 if i < u else 0 for i in range(n)]

    @property
    def nsets(self):
        return (len(PERM_COLUMNS) + CHUNK - 1) // CHUNK

    def products(self):
        """Permutation grand products (permutation/prover.rs:355-515)."""
        n, u = self.n, self.u
        beta, gamma = self.beta, self.gamma
        last_z = 1
        deltaomega_col = 0          # column index of the running delta power
        blinds = iter(self.blinds)
        zs = []
        for s in range(self.nsets):
            cols = list(range(s * CHUNK, min((s + 1) * CHUNK, len(PERM_COLUMNS))))
            modified = [1] * n
            for j in cols:
                v = self.cols[j]
                sig = self.sigma[j]
                for i in range(n):
                    modified[i] = modified[i] * (beta * sig[i] + gamma + v[i]) % R
            modified = [pow(m, R - 2, R) for m in modified]
            for j in cols:
                v = self.cols[j]
                dj = self.delta_pows[deltaomega_col]
                for i in range(n):
                    modified[i] = modified[i] * (dj * self.omega_pows[i] * beta + gamma + v[i]) % R
                deltaomega_col += 1
            z = [last_z]
            for i in range(1, n):
                z.append(z[i - 1] * modified[i - 1] % R)
            for i in range(n - BLINDING, n):
                z[i] = next(blinds)
            last_z = z[u]
            zs.append(z)
        return zs

    # ---- numerator inputs -------------------------------------------------

    def custom_graph(self):
        return gate_graph()

    def base_inputs(self):
        """evaluate_numerator inputs on the base domain (LagrangeCoeff):
        rot_scale 1, omega, g_coset 1."""
        n = self.n
        flat = lambda cs: mont([x for c in cs for x in c])
        return dict(
            n=n, fixed=flat(self.cols[5:]), advice=flat(self.cols[:5]),
            beta=mont([self.beta])[0], gamma=mont([self.gamma])[0], theta=mont([self.theta])[0],
            trash_challenge=mont([self.trash])[0], y=mont([self.y])[0], omega=mont([self.omega])[0],
            g_coset=mont([1])[0], rot_scale=1, last_rotation=-(BLINDING + 1), chunk_len=CHUNK,
            l0=mont(self.l0), l_last=mont(self.l_last), l_active=mont(self.l_active),
            permutation_products=flat(self.z), permutation_columns=PERM_COLUMNS,
            permutation_pk_cosets=flat(self.sigma), custom_graph=self.custom_graph())

    def extended_inputs(self, threads=None):
        """The same inputs on the extended coset (ExtendedLagrangeCoeff):
        every column Lagrange -> coefficients -> coset evaluations."""
        k = self.k
        dom, t_evals = O.domain(DEGREE, k)
        self.domain, self.t_evals = dom, t_evals
        N = 1 << dom.extended_k

        def ext(col):
            coeffs = O.ifft(mont(col), k, threads)
            return O.coeff_to_extended(coeffs, dom, threads)

        def flat(cs):
            return np.concatenate([ext(c) for c in cs])

        base = self.base_inputs()
        base.update(
            n=N, fixed=flat(self.cols[5:]), advice=flat(self.cols[:5]),
            omega=np.array(dom.fr("extended_omega")), g_coset=np.array(dom.fr("g_coset")),
            rot_scale=1 << (dom.extended_k - k),
            l0=ext(self.l0), l_last=ext(self.l_last), l_active=ext(self.l_active),
            permutation_products=flat(self.z), permutation_pk_cosets=flat(self.sigma))
        return base


class FastCircuit:
    """The same circuit family at sizes where Python integers are too slow
    (k up to 20): Montgomery numpy columns from the oracle's seeded RNG, a
    satisfying witness d = a*b + c on the usable rows, and copy cycles of
    two that make a quarter of the usable rows of e equal to rows of a.
    Provides what the GPU key and prover take (fixed, advice, mapping)."""

    def __init__(self, k, seed=1):
        self.k = k
        n = self.n = 1 << k
        u = self.u = n - BLINDING - 1
        vals = O.random_fr(7 * n, seed)
        a, b, c, d, e, s1, s2 = (vals[i * n:(i + 1) * n].copy() for i in range(7))
        d[:u] = O.mul_add(a[:u], b[:u], c[:u])
        one = mont([1])[0]
        q = np.zeros((n, 4), dtype=np.uint64)
        q[:u] = one
        s1[u:] = 0
        s2[u:] = 0
        m = np.zeros((len(PERM_COLUMNS), n, 2), dtype=np.uint32)
        m[:, :, 0] = np.arange(len(PERM_COLUMNS), dtype=np.uint32)[:, None]
        m[:, :, 1] = np.arange(n, dtype=np.uint32)[None, :]
        rows = np.random.default_rng(seed).permutation(u)
        pairs = u // 4
        ei, aj = rows[:pairs], rows[pairs:2 * pairs]
        e[ei] = a[aj]
        m[4, ei, 0], m[4, ei, 1] = 0, aj
        m[0, aj, 0], m[0, aj, 1] = 4, ei
        self.fixed = np.concatenate([q, s1, s2])
        self.advice = np.concatenate([a, b, c, d, e])
        self.mapping = m.reshape(-1)
        self.omega = omega_int(k)
        self.delta = fr_const("delta")
        self.delta_pows = [pow(self.delta, j, R) for j in range(len(PERM_COLUMNS))]
        # what the CPU reference prover's key reads, as Montgomery arrays
        self.cols = [a, b, c, d, e, q, s1, s2]
        one = mont([1])[0]
        wpow = O.shifted_prefix_product(np.tile(mont([self.omega])[0], (n, 1)), one)
        dpow = mont(self.delta_pows)
        zero = np.zeros((n, 4), dtype=np.uint64)
        self.sigma = [O.mul_add(dpow[m[j, :, 0]], wpow[m[j, :, 1]], zero) for j in range(len(PERM_COLUMNS))]
        lag = np.zeros((3, n, 4), dtype=np.uint64)
        lag[0, 0] = one
        lag[1, u] = one
        lag[2, :u] = one
        self.l0, self.l_last, self.l_active = lag[0], lag[1], lag[2]

    @property
    def nsets(self):
        return (len(PERM_COLUMNS) + CHUNK - 1) // CHUNK

    def custom_graph(self):
        return gate_graph()


def random_inputs(n, seed, *, rot_scale=4, sets=3, graph=None):
    """Unstructured numerator inputs of the synthetic circuit's shape
    (5 advice, 3 fixed, 8 permutation columns): for parity checks at sizes
    where building a satisfying witness in Python is too slow."""
    ncols = len(PERM_COLUMNS)
    vals = O.random_fr(n * (ADVICE + FIXED + ncols + sets + 3) + 8, seed)
    off = [0]

    def take(m):
        a = vals[off[0]:off[0] + m]
        off[0] += m
        return a

    adv, fix, sig, zs = take(ADVICE * n), take(FIXED * n), take(ncols * n), take(sets * n)
    l0, ll, la = take(n), take(n), take(n)
    sc = take(8)
    return dict(
        n=n, fixed=fix, advice=adv, beta=sc[0], gamma=sc[1], theta=sc[2], trash_challenge=sc[3], y=sc[4],
        omega=sc[5], g_coset=sc[6], rot_scale=rot_scale, last_rotation=-(BLINDING + 1), chunk_len=CHUNK,
        l0=l0, l_last=ll, l_active=la, permutation_products=zs, permutation_columns=PERM_COLUMNS,
        permutation_pk_cosets=sig,
        custom_graph=graph or G.RawGraphBuffers(G.custom_gates_graph([gate_expression()])))


DESCRIPTION = (b"synthetic PLONK circuit: advice a b c d e, fixed q s1 s2, gate q*(a*b+c-d), "
               b"copy permutation over a b c d e q s1 s2, degree 5")
ADVICE_QUERIES = [(c, 0) for c in range(ADVICE)]
FIXED_QUERIES = [(c, 0) for c in range(FIXED)]


def mapping(c):
    """(j', i') of the next cell of every permuted cell's cycle, as the
    keygen's flat uint32 array [2 * (j * n + i)]."""
    n = c.n
    m = np.zeros((len(PERM_COLUMNS), n, 2), dtype=np.uint32)
    m[:, :, 0] = np.arange(len(PERM_COLUMNS), dtype=np.uint32)[:, None]
    m[:, :, 1] = np.arange(n, dtype=np.uint32)[None, :]
    for cyc in c.cycles:
        for t, (j, i) in enumerate(cyc):
            jj, ii = cyc[(t + 1) % len(cyc)]
            m[j, i] = (jj, ii)
    return m.reshape(-1)


# ---------------------------------------------------------------------------
# The family with every argument of the prover: lookups, a trashcan and an
# instance column next to the gate and the copy permutation.

ADVICE_X = 7       # a b c d e f g
FIXED_X = 7        # q s1 s2 t1 t2 qp qt
INSTANCE_X = 1     # p
PERM_COLUMNS_X = PERM_COLUMNS + [(2, 0)]
ADVICE_QUERIES_X = [(c, 0) for c in range(ADVICE_X)]
FIXED_QUERIES_X = [(c, 0) for c in range(FIXED_X)]
INSTANCE_QUERIES_X = [(0, 0)]
DESCRIPTION_X = (b"synthetic PLONK circuit with arguments: advice a b c d e f g, fixed q s1 s2 t1 t2 qp qt, "
                 b"instance p, gates q*(a*b+c-d) and qp*(a-p), lookups (f, g) in (t1, t2) and e in s1, "
                 b"trashcan qt: (d-a*b-c, e*(d-a*b-c)), copy permutation over a b c d e q s1 s2 p, degree 5")


def _adv(i):
    return ("advice", i, 0)


def _fix(i):
    return ("fixed", i, 0)


def _sub(x, y):
    return ("sum", x, ("neg", y))


def gate_expressions_x():
    a, b, c, d = (_adv(i) for i in range(4))
    return [("prod", _fix(0), ("sum", ("sum", ("prod", a, b), c), ("neg", d))),
            ("prod", _fix(5), _sub(a, ("instance", 0, 0)))]


def lookup_expressions_x():
    """(input expressions, table expressions) per lookup"""
    return [([_adv(5), _adv(6)], [_fix(3), _fix(4)]),
            ([("prod", _fix(0), _adv(4))], [_fix(1)])]


def trash_expressions_x():
    """(selector fixed column, constraint expressions) per trashcan"""
    r = _sub(_sub(_adv(3), ("prod", _adv(0), _adv(1))), _adv(2))
    return [(6, [r, ("prod", _adv(4), r)])]


class ArgCircuit(Circuit):
    """Circuit plus: advice f, g holding (t1, t2) table rows (lookup 0,
    inputs (f, g) in table (t1, t2)); lookup 1, q * e in table s1 (e's free
    cells copy s1 values); a trashcan with selector qt on odd usable rows
    whose constraints vanish there; an instance column p of `m` public values
    equal to a's first rows, tied to them by the gate qp * (a - p) and by
    copy constraints (the instance column is the ninth permuted column);
    with committed = 1 that column is a committed instance
    (nb_committed_instances of create_proof)."""

    def __init__(self, k, seed=1, m=6, committed=0):
        super().__init__(k, seed, reserved={(0, r) for r in range(m)})
        self.committed = committed      # leading instance columns that are committed
        rnd, n, u = self.rnd, self.n, self.u
        rv = lambda: rnd.randrange(R)
        a, b, c, d, e, q, s1, s2 = self.cols
        # lookup 1 (q * e in s1): every usable row of e holds an s1 value.
        # e's cells copied from a..d in cycles hold those values instead, so
        # s1 takes them at the same rows (s1 itself is in cycles too: its
        # copied cells already equal their source)
        in_cycle = {cell for cyc in self.cycles for cell in cyc}
        taken = set()     # rows of s1 whose value an e cell needs
        for i in range(u):
            if (4, i) not in in_cycle:
                continue
            if (6, i) not in in_cycle and i not in taken:
                j = i
            else:
                j = next(r for r in range(u) if (6, r) not in in_cycle and r not in taken)
            s1[j] = e[i]
            taken.add(j)
        for i in range(u):
            if (4, i) not in in_cycle:
                e[i] = s1[rnd.randrange(u)]
        t1 = [rv() if i < u else 0 for i in range(n)]
        t2 = [rv() if i < u else 0 for i in range(n)]
        rows = [rnd.randrange(u) for _ in range(n)]
        f = [t1[rows[i]] if i < u else rv() for i in range(n)]
        g = [t2[rows[i]] if i < u else rv() for i in range(n)]
        qp = [1 if i < m else 0 for i in range(n)]
        qt = [1 if (i < u and i % 2 == 1) else 0 for i in range(n)]
        self.instance = [[a[i] for i in range(m)]]
        self.cols = [a, b, c, d, e, q, s1, s2]
        self.advice_values = [a, b, c, d, e, f, g]
        self.fixed_values = [q, s1, s2, t1, t2, qp, qt]
        for r in range(m):
            self.cycles.append([(0, r), (8, r)])
        # sigma over the nine permuted columns; cells outside cycles map to themselves
        mapping = {}
        for cyc in self.cycles:
            for t, cell in enumerate(cyc):
                mapping[cell] = cyc[(t + 1) % len(cyc)]
        ncol = len(PERM_COLUMNS_X)
        self.delta_pows = [pow(self.delta, j, R) for j in range(ncol)]
        self.sigma = [[self.delta_pows[mapping.get((j, i), (j, i))[0]] *
                       self.omega_pows[mapping.get((j, i), (j, i))[1]] % R for i in range(n)]
                      for j in range(ncol)]

    def custom_graph(self):
        return G.RawGraphBuffers(G.custom_gates_graph(gate_expressions_x()))

    def lookup_graphs(self):
        """(input compression, table compression, numerator graph) per lookup"""
        return [(G.RawGraphBuffers(G.compression_graph(i)), G.RawGraphBuffers(G.compression_graph(t)),
                 G.RawGraphBuffers(G.lookup_graph(i, t))) for i, t in lookup_expressions_x()]

    def trash_graphs(self):
        """(graph, selector fixed column) per trashcan"""
        return [(G.RawGraphBuffers(G.trash_graph(cs)), sel) for sel, cs in trash_expressions_x()]


def mapping_x(c):
    """mapping() over the nine permuted columns of ArgCircuit"""
    n = c.n
    m = np.zeros((len(PERM_COLUMNS_X), n, 2), dtype=np.uint32)
    m[:, :, 0] = np.arange(len(PERM_COLUMNS_X), dtype=np.uint32)[:, None]
    m[:, :, 1] = np.arange(n, dtype=np.uint32)[None, :]
    for cyc in c.cycles:
        for t, (j, i) in enumerate(cyc):
            jj, ii = cyc[(t + 1) % len(cyc)]
            m[j, i] = (jj, ii)
    return m.reshape(-1)
