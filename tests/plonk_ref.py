"""Reference-algorithm CPU prover and a designated verifier for the
synthetic circuit family.  TEST INFRASTRUCTURE ONLY.

cpu_keygen / cpu_prove restate the reference prover's protocol
(vendor/midnight-proofs-0.8.1/src/plonk/prover.rs:68-575, :769-1010,
:1931-2080; permutation/prover.rs:327-712; vanishing/prover.rs:41-406;
poly/kzg/mod.rs:1978-2570; transcript/implementors.rs:13-33) on the CPU
oracle's primitives (Pippenger MSM, best_fft, coset FFT, numerator, grand
products, kate division).  The GPU pipeline must produce the same bytes.

verify() is a designated verifier: it knows the SRS trapdoor tau, so the
KZG pairing check e(pi, [tau]G2) = e(C - vG + x3 pi, G2) of the multi-open
(poly/kzg/mod.rs:1806-1976 multi_prepare) becomes the G1 equation
[tau] pi = C - vG + x3 pi.  It replays the transcript, recomputes the
numerator at x from the proof's evaluations (plonk/verifier.rs; gate,
permutation/verifier.rs and vanishing/verifier.rs expressions) and checks
the opening, so it pins the protocol's algebra independently of the
provers.
"""
import hashlib

import numpy as np

import numerator_graph as G
import oracle as O
import synth_circuit as S

R = G.FR_MODULUS
KEY = b"Domain separator for transcript"


def to_int(a):
    return S.ints(np.asarray(a, dtype=np.uint64).reshape(1, 4))[0]


def to_fr(v):
    return S.mont([v % R])[0]


def fr_bytes(f):
    return to_int(f).to_bytes(32, "little")


def jac_of_affine(a):
    out = O.g1j()
    O.lib().orc_g1j_from_affine(O.P(out), O.P(np.ascontiguousarray(a, dtype=np.uint64)))
    return out


def g_add(a, b):
    out = O.g1j()
    O.lib().orc_g1j_add(O.P(out), O.P(np.ascontiguousarray(a)), O.P(np.ascontiguousarray(b)))
    return out


def g_mul(p, s_int):
    out = O.g1j()
    O.lib().orc_g1j_mul(O.P(out), O.P(np.ascontiguousarray(p)), O.P(to_fr(s_int)))
    return out


def g_neg(p):
    out = O.g1j()
    O.lib().orc_g1j_neg(O.P(out), O.P(np.ascontiguousarray(p)))
    return out


class Transcript:
    def __init__(self, proof=b""):
        self.h = hashlib.blake2b(digest_size=64, key=KEY)
        self.proof = bytearray()
        self.rd = memoryview(bytes(proof))
        self.pos = 0

    def absorb(self, b):
        self.h.update(b"\x01" + bytes(b))

    def common_scalar(self, f_int):
        self.absorb(f_int.to_bytes(32, "little"))

    def common_point(self, jac):
        self.absorb(bytes(O.compress(O.to_affine(jac))))

    def write_scalar(self, f_int):
        b = f_int.to_bytes(32, "little")
        self.absorb(b)
        self.proof += b

    def write_point(self, jac):
        b = O.compress(O.to_affine(jac))
        self.absorb(b)
        self.proof += b

    def squeeze(self):
        self.h.update(b"\x00")
        return int.from_bytes(self.h.copy().digest(), "little") % R

    def read_point(self):
        b = bytes(self.rd[self.pos:self.pos + 48])
        if len(b) != 48:
            raise ValueError("proof too short")
        self.pos += 48
        self.absorb(b)
        return jac_of_affine(O.decompress(b))

    def read_scalar(self):
        b = bytes(self.rd[self.pos:self.pos + 32])
        if len(b) != 32:
            raise ValueError("proof too short")
        self.pos += 32
        v = int.from_bytes(b, "little")
        if v >= R:
            raise ValueError("non-canonical scalar")
        self.absorb(b)
        return v


# ------------------------------------------------------------------ keygen

class Spec:
    """The constraint system of a circuit family: column counts, queries,
    gate, lookup and trash expressions, permuted columns, key description."""

    def __init__(self, args=False, committed=0):
        self.committed = committed
        if args:
            self.advice, self.fixed, self.instance = S.ADVICE_X, S.FIXED_X, S.INSTANCE_X
            self.perm = S.PERM_COLUMNS_X
            self.aq, self.fq, self.iq = S.ADVICE_QUERIES_X, S.FIXED_QUERIES_X, S.INSTANCE_QUERIES_X
            self.gates = S.gate_expressions_x()
            self.lookups = S.lookup_expressions_x()
            self.trash = S.trash_expressions_x()
            self.description = S.DESCRIPTION_X
        else:
            self.advice, self.fixed, self.instance = S.ADVICE, S.FIXED, 0
            self.perm = S.PERM_COLUMNS
            self.aq, self.fq, self.iq = S.ADVICE_QUERIES, S.FIXED_QUERIES, []
            self.gates = [S.gate_expression()]
            self.lookups, self.trash = [], []
            self.description = S.DESCRIPTION

    @staticmethod
    def of(circuit):
        return Spec(isinstance(circuit, S.ArgCircuit), getattr(circuit, "committed", 0))


class CpuKey:
    def __init__(self, circuit, tau_int, threads=None):
        c = circuit
        self.c = c
        self.k, self.n = c.k, c.n
        self.tau = tau_int
        n, k = self.n, self.k
        self.srs = O.srs(to_fr(tau_int), n, threads)
        self.nsets = c.nsets
        self.dom, self.t_evals = O.domain(S.DEGREE, k)
        self.N = 1 << self.dom.extended_k
        self.bf, self.u = S.BLINDING, c.u
        self.pieces = S.DEGREE - 1
        self.omega = c.omega
        self.zeta = S.fr_const("zeta")
        self.delta = S.fr_const("delta")
        m = lambda col: col if isinstance(col, np.ndarray) else S.mont(col)
        self.spec = Spec.of(c)
        self.fixed_lag = [m(col) for col in getattr(c, "fixed_values", c.cols[5:])]
        self.sigma_lag = [m(col) for col in c.sigma]
        self.fixed_coeff = [O.ifft(v, k, threads) for v in self.fixed_lag]
        self.sigma_coeff = [O.ifft(v, k, threads) for v in self.sigma_lag]
        self.fixed_com = [O.msm(self.srs, p, threads) for p in self.fixed_coeff]
        self.sigma_com = [O.msm(self.srs, p, threads) for p in self.sigma_coeff]
        ext = lambda v: O.coeff_to_extended(O.ifft(v, k, threads), self.dom, threads)
        self.fixed_ext = np.concatenate([O.coeff_to_extended(p, self.dom, threads) for p in self.fixed_coeff])
        self.sigma_ext = np.concatenate([O.coeff_to_extended(p, self.dom, threads) for p in self.sigma_coeff])
        self.l0_ext, self.l_last_ext, self.l_active_ext = (ext(m(v)) for v in (c.l0, c.l_last, c.l_active))
        hv = hashlib.blake2b(digest_size=64, person=b"Halo2-Verify-Key")
        hv.update(bytes([3, k]))
        for coms in (self.fixed_com, self.sigma_com):
            hv.update(len(coms).to_bytes(4, "little"))
            for p in coms:
                hv.update(O.to_affine(p).astype("<u8").tobytes())
        hv.update(self.spec.description)
        self.repr = int.from_bytes(hv.digest(), "little") % R
        sp = self.spec
        self.custom_graph = G.RawGraphBuffers(G.custom_gates_graph(sp.gates))
        self.compress_graphs = [(G.RawGraphBuffers(G.compression_graph(i)), G.RawGraphBuffers(G.compression_graph(t)))
                                for i, t in sp.lookups]
        self.lookup_graphs = [G.RawGraphBuffers(G.lookup_graph(i, t)) for i, t in sp.lookups]
        self.trash_graphs = [G.RawGraphBuffers(G.trash_graph(cs)) for _, cs in sp.trash]


class VerifierKey:
    """What the designated verifier reads of a key, without the prover's
    polynomials: sizes, domain constants and the SRS trapdoor (for the GPU
    keys of circuits too large for CpuKey)."""

    def __init__(self, k, tau_int, args=False, committed=0):
        self.k, self.n = k, 1 << k
        self.spec = Spec(args, committed)
        self.tau = tau_int
        self.bf = S.BLINDING
        self.u = self.n - self.bf - 1
        self.pieces = S.DEGREE - 1
        self.omega = S.omega_int(k)
        self.delta = S.fr_const("delta")
        self.nsets = (len(self.spec.perm) + S.CHUNK - 1) // S.CHUNK


# ------------------------------------------------------------------ prover

def _rot(key, x, rot):
    w = key.omega if rot >= 0 else pow(key.omega, R - 2, R)
    return x * pow(w, abs(rot), R) % R


def _lincomb(polys, scalars, n):
    """sum_j scalars[j] polys[j] (scalars as integers) by the oracle"""
    return O.lincomb(np.concatenate([np.asarray(p).reshape(n, 4) for p in polys]),
                     S.mont([s % R for s in scalars]), n)


def _eval(p, x_int):
    return to_int(O.poly_eval(p, to_fr(x_int)))


def parallelize_chunks(n, threads):
    """(offset, length) of the chunks utils/arithmetic.rs:138-190
    parallelize cuts n rows into for `threads` workers: n % threads chunks of
    n // threads + 1 rows, then chunks of n // threads rows (none when that
    is 0), in row order."""
    base, cutoff = divmod(n, threads)
    out = [(c * (base + 1), base + 1) for c in range(cutoff)]
    if base:
        out += [(cutoff * (base + 1) + c * base, base) for c in range(threads - cutoff)]
    return out


def sample_random_seed_offsets(n, threads):
    """The offsets vanishing/prover.rs:50-66 pairs with its seeds, in the
    order the seeds are drawn: (0..).step_by(chunk + 1).take(n % threads)
    chained with (n % threads) * (chunk + 1) stepping by chunk when chunk != 0,
    at most `threads` of them (zip stops drawing seeds with the offsets)."""
    chunk = n // threads
    first = [i * (chunk + 1) for i in range(n % threads)]
    rest = [(n % threads) * (chunk + 1) + i * chunk for i in range(threads)] if chunk != 0 else []
    return (first + rest)[:threads]


def sample_random(rng, n, rng_threads):
    """vanishing::Argument::sample_random: one 32-byte seed per chunk drawn
    from the prover's RNG, each chunk filled from its own ChaCha20Rng."""
    chunks = parallelize_chunks(n, rng_threads)
    seeds = [rng.fill_bytes(32) for _ in sample_random_seed_offsets(n, rng_threads)]
    assert len(seeds) == len(chunks)
    return np.concatenate([O.ChaCha(seed_bytes=sd).random_fr(ln) for sd, (_, ln) in zip(seeds, chunks)])


def _perm_columns(key, adv, inst_lag):
    """the permuted columns' Lagrange values, in permutation order"""
    n = key.n
    out = []
    for kind, idx in key.spec.perm:
        if kind == 0:
            out.append(key.fixed_lag[idx])
        elif kind == 1:
            out.append(adv[idx * n:(idx + 1) * n])
        else:
            out.append(inst_lag[idx * n:(idx + 1) * n])
    return np.concatenate(out)


def cpu_prove(key, advice_cols, seed, threads=None, rng_threads=1, instance=None):
    """advice_cols: list of advice columns of n ints (Lagrange values), or
    their Montgomery limbs column after column; instance: list of instance
    columns (ints) of the key's instance count; rng_threads: the thread count
    of the reference prover whose random draws to follow."""
    c, n, k = key.c, key.n, key.k
    sp = key.spec
    A, L, TR, I = sp.advice, len(sp.lookups), len(sp.trash), sp.instance
    Sn = (len(sp.perm) + S.CHUNK - 1) // S.CHUNK
    u, bf = key.u, key.bf
    rng = O.ChaCha(seed=seed)
    if isinstance(advice_cols, np.ndarray):     # Montgomery limbs, column after column
        adv = np.array(advice_cols, dtype=np.uint64, copy=True).reshape(A * n, 4)
    else:
        adv = S.mont([v for col in advice_cols for v in col]).copy()
    for col in range(A):
        adv[col * n + u:(col + 1) * n] = rng.random_fr(n - u)
    lk_blinds = [(rng.random_fr(n - u), rng.random_fr(n - u)) for _ in range(L)]
    perm_blinds = rng.random_fr(Sn * bf)
    lk_prod_blinds = rng.random_fr(L * bf) if L else None
    rand_poly = sample_random(rng, n, rng_threads)
    limb_blinds = rng.random_fr(key.pieces - 1)

    T = Transcript()
    T.common_scalar(key.repr)
    # instance columns (compute_instances, prover.rs:1255-1286)
    instance = instance if instance is not None else [[] for _ in range(I)]
    assert len(instance) == I
    inst_lag, inst_coeff = None, []
    if I:
        vals = []
        for col in instance:
            assert len(col) <= u, "instance column longer than the usable rows"
            vals += [v % R for v in col] + [0] * (n - len(col))
        inst_lag = S.mont(vals)
        inst_coeff = [O.ifft(inst_lag[i * n:(i + 1) * n], k, threads) for i in range(I)]
        for i, col in enumerate(instance):
            if i < sp.committed:
                T.common_point(O.msm(key.srs, inst_coeff[i], threads))
            else:
                T.common_scalar(len(col))
                for v in col:
                    T.common_scalar(v % R)
    fixed_all = np.concatenate(key.fixed_lag)
    adv_coeff = [O.ifft(adv[i * n:(i + 1) * n], k, threads) for i in range(A)]
    for p in adv_coeff:
        T.write_point(O.msm(key.srs, p, threads))
    theta = T.squeeze()

    # lookups (lookup/prover.rs:737-1072, permute_expression_pair :1573-1690)
    ci_all, ct_all, pin_all, ptab_all = [], [], [], []
    for l in range(L):
        gi, gt = key.compress_graphs[l]
        ci = O.graph_rows(n, gi, theta=to_fr(theta), fixed=fixed_all, advice=adv, instance=inst_lag,
                          threads=threads)
        ct = O.graph_rows(n, gt, theta=to_fr(theta), fixed=fixed_all, advice=adv, instance=inst_lag,
                          threads=threads)
        perm = O.lookup_permute(ci, ct, u)
        if perm is None:
            raise ValueError("lookup input not in its table")
        ci_all.append(ci)
        ct_all.append(ct)
        pin_all.append(np.concatenate([perm[0], lk_blinds[l][0]]))
        ptab_all.append(np.concatenate([perm[1], lk_blinds[l][1]]))
    lk_in_coeff = [O.ifft(v, k, threads) for v in pin_all]
    lk_tab_coeff = [O.ifft(v, k, threads) for v in ptab_all]
    for l in range(L):
        T.write_point(O.msm(key.srs, lk_in_coeff[l], threads))
        T.write_point(O.msm(key.srs, lk_tab_coeff[l], threads))
    beta, gamma = T.squeeze(), T.squeeze()

    z_lag = O.permutation_products(_perm_columns(key, adv, inst_lag), np.concatenate(key.sigma_lag), n, S.CHUNK,
                                   to_fr(beta), to_fr(gamma), to_fr(key.omega),
                                   S.mont([pow(key.delta, j, R) for j in range(len(sp.perm))]), perm_blinds, bf,
                                   threads)
    z_coeff = [O.ifft(z_lag[s * n:(s + 1) * n], k, threads) for s in range(Sn)]
    lz_coeff = []
    if L:
        lz = O.lookup_products(np.concatenate(ci_all), np.concatenate(ct_all), np.concatenate(pin_all),
                               np.concatenate(ptab_all), n, L, to_fr(beta), to_fr(gamma), lk_prod_blinds, bf)
        lz_coeff = [O.ifft(lz[l * n:(l + 1) * n], k, threads) for l in range(L)]
    for p in z_coeff + lz_coeff:
        T.write_point(O.msm(key.srs, p, threads))
    trash = T.squeeze()
    tr_coeff = []
    for t in range(TR):
        tv = O.graph_rows(n, key.trash_graphs[t], trash_challenge=to_fr(trash), fixed=fixed_all, advice=adv,
                          instance=inst_lag, threads=threads)
        tr_coeff.append(O.ifft(tv, k, threads))
        T.write_point(O.msm(key.srs, tr_coeff[-1], threads))
    T.write_point(O.msm(key.srs, rand_poly, threads))
    y = T.squeeze()

    N = key.N
    ext = lambda ps: np.concatenate([O.coeff_to_extended(p, key.dom, threads) for p in ps]) if ps else None
    nu = O.numerator(N, fixed=key.fixed_ext, advice=ext(adv_coeff),
                     instance=ext(inst_coeff),
                     beta=to_fr(beta), gamma=to_fr(gamma),
                     theta=to_fr(theta), trash_challenge=to_fr(trash), y=to_fr(y),
                     omega=np.array(key.dom.fr("extended_omega")), g_coset=np.array(key.dom.fr("g_coset")),
                     rot_scale=N // n, last_rotation=-(bf + 1), chunk_len=S.CHUNK, l0=key.l0_ext,
                     l_last=key.l_last_ext, l_active=key.l_active_ext, permutation_products=ext(z_coeff),
                     permutation_columns=sp.perm, permutation_pk_cosets=key.sigma_ext,
                     lookup_product_cosets=ext(lz_coeff), lookup_input_cosets=ext(lk_in_coeff),
                     lookup_table_cosets=ext(lk_tab_coeff), lookup_graphs=key.lookup_graphs,
                     trash_cosets=ext(tr_coeff), trash_selector_fixed_indices=[sel for sel, _ in sp.trash],
                     trash_graphs=key.trash_graphs, custom_graph=key.custom_graph, threads=threads)
    # h truncated to (n - 1) pieces of n - 1 coefficients, each padded to n with
    # the blinding telescope (vanishing/prover.rs:155-182, 304-318)
    h = O.extended_to_coeff(O.mul_repeating(nu, key.t_evals), key.dom, threads)
    pieces = np.zeros((key.pieces * n, 4), dtype=np.uint64)
    for i in range(key.pieces):
        pieces[i * n:i * n + n - 1] = h[i * (n - 1):(i + 1) * (n - 1)]
    for i in range(1, key.pieces):
        pieces[(i - 1) * n + n - 1] = limb_blinds[i - 1]
        pieces[i * n] = O.fr_op("sub", pieces[i * n], limb_blinds[i - 1])
    h_pieces = [pieces[i * n:(i + 1) * n] for i in range(key.pieces)]
    for p in h_pieces:
        T.write_point(O.msm(key.srs, p, threads))
    x = T.squeeze()

    committed_q = [(col, rot) for col, rot in sp.iq if col < sp.committed]
    for col, rot in committed_q:
        T.write_scalar(_eval(inst_coeff[col], _rot(key, x, rot)))
    for col, rot in sp.aq:
        T.write_scalar(_eval(adv_coeff[col], _rot(key, x, rot)))
    for col, rot in sp.fq:
        T.write_scalar(_eval(key.fixed_coeff[col], _rot(key, x, rot)))
    T.write_scalar(_eval(rand_poly, x))
    for p in key.sigma_coeff:
        T.write_scalar(_eval(p, x))
    x_next, x_last, x_prev = _rot(key, x, 1), _rot(key, x, -(bf + 1)), _rot(key, x, -1)
    for s in range(Sn):
        T.write_scalar(_eval(z_coeff[s], x))
        T.write_scalar(_eval(z_coeff[s], x_next))
        if s + 1 < Sn:
            T.write_scalar(_eval(z_coeff[s], x_last))
    for l in range(L):      # lookup/prover.rs:1341-1525
        T.write_scalar(_eval(lz_coeff[l], x))
        T.write_scalar(_eval(lz_coeff[l], x_next))
        T.write_scalar(_eval(lk_in_coeff[l], x))
        T.write_scalar(_eval(lk_in_coeff[l], x_prev))
        T.write_scalar(_eval(lk_tab_coeff[l], x))
    for t in range(TR):
        T.write_scalar(_eval(tr_coeff[t], x))
    sf = pow(x, n - 1, R)
    h_poly = _lincomb(h_pieces, [pow(sf, i, R) for i in range(key.pieces)], n)

    queries = [(("adv", col), adv_coeff[col], _rot(key, x, rot)) for col, rot in sp.aq]
    queries += [(("inst", col), inst_coeff[col], _rot(key, x, rot)) for col, rot in committed_q]
    for s in range(Sn):
        queries += [(("z", s), z_coeff[s], x), (("z", s), z_coeff[s], x_next)]
    for s in range(Sn - 2, -1, -1):
        queries.append((("z", s), z_coeff[s], x_last))
    for l in range(L):      # lookup/prover.rs:1527-1572
        queries += [(("lz", l), lz_coeff[l], x), (("li", l), lk_in_coeff[l], x), (("lt", l), lk_tab_coeff[l], x),
                    (("li", l), lk_in_coeff[l], x_prev), (("lz", l), lz_coeff[l], x_next)]
    queries += [(("trash", t), tr_coeff[t], x) for t in range(TR)]
    queries += [(("fixed", col), key.fixed_coeff[col], _rot(key, x, rot)) for col, rot in sp.fq]
    queries += [(("sigma", j), p, x) for j, p in enumerate(key.sigma_coeff)]
    queries += [(("h",), h_poly, x), (("rand",), rand_poly, x)]
    _multi_open_prove(T, key, queries, n, threads)
    return bytes(T.proof)


def _intermediate_sets(queries):
    """construct_intermediate_set_metadata (poly/kzg/utils.rs:47-160)."""
    points, cids, cpoints, cobj = [], [], [], []
    for cid, obj, pt in queries:
        if pt not in points:
            points.append(pt)
        pi = points.index(pt)
        if cid not in cids:
            cids.append(cid)
            cpoints.append([])
            cobj.append(obj)
        ci = cids.index(cid)
        assert pi not in cpoints[ci], "duplicated query"
        cpoints[ci].append(pi)
    set_index = {}
    cset = []
    for ps in cpoints:
        key = frozenset(ps)
        if key not in set_index:
            set_index[key] = len(set_index)
        cset.append(set_index[key])
    set_points = [None] * len(set_index)
    for key, i in set_index.items():
        set_points[i] = sorted(key)
    return points, cids, cpoints, cobj, cset, set_points


def _multi_open_prove(T, key, queries, n, threads):
    x1, x2 = T.squeeze(), T.squeeze()
    points, cids, cpoints, cobj, cset, set_points = _intermediate_sets(queries)
    nsets = len(set_points)
    qpolys = []
    for s in range(nsets):
        members = [cobj[c] for c in range(len(cids)) if cset[c] == s]
        qpolys.append(_lincomb(members, [pow(x1, j, R) for j in range(len(members))], n))
    order = sorted(range(nsets), key=lambda i: (len(set_points[i]), i))
    divided = []
    for s in order:
        p = qpolys[s]
        for pi in set_points[s]:
            p = O.kate_division(p, to_fr(points[pi]))
        divided.append(p)
    f = _lincomb(divided, [pow(x2, i, R) for i in range(nsets)], n)
    T.write_point(O.msm(key.srs, f, threads))
    x3 = T.squeeze()
    for s in order:
        T.write_scalar(_eval(qpolys[s], x3))
    x4 = T.squeeze()
    final = _lincomb([qpolys[s] for s in order] + [f], [pow(x4, i, R) for i in range(nsets + 1)], n)
    v = _eval(final, x3)
    fi = S.ints(final)
    fi[0] = (fi[0] - v) % R
    pi_poly = O.kate_division(S.mont(fi), to_fr(x3))
    T.write_point(O.msm(key.srs, pi_poly, threads))


# ------------------------------------------------------------------ verifier

def _lagrange_at(key, x, i):
    """l_i(x) = omega^i (x^n - 1) / (n (x - omega^i))"""
    n = key.n
    wi = pow(key.omega, i, R)
    return wi * (pow(x, n, R) - 1) * pow(n * (x - wi) % R, R - 2, R) % R


def _expr_at(e, ev, key):
    """an expression tree (numerator_graph's tags) at x from the evaluations:
    ev(kind, column, rotation)"""
    tag = e[0]
    if tag == "constant":
        return e[1] % R
    if tag in ("fixed", "advice", "instance"):
        return ev(tag, e[1], e[2])
    if tag == "neg":
        return -_expr_at(e[1], ev, key) % R
    if tag == "sum":
        return (_expr_at(e[1], ev, key) + _expr_at(e[2], ev, key)) % R
    if tag == "prod":
        return _expr_at(e[1], ev, key) * _expr_at(e[2], ev, key) % R
    if tag == "scaled":
        return _expr_at(e[1], ev, key) * e[2] % R
    raise ValueError("unknown expression tag %r" % (tag,))


def verify(vk, proof, key, instance=None):
    """vk: (fixed commitments, permutation commitments, transcript repr int);
    instance: the proof's instance columns (ints).  Returns True when the
    proof verifies."""
    fixed_com, sigma_com, repr_int = vk
    sp = key.spec
    n, A, bf = key.n, sp.advice, key.bf
    L, TR, I = len(sp.lookups), len(sp.trash), sp.instance
    Sn = (len(sp.perm) + S.CHUNK - 1) // S.CHUNK
    instance = instance if instance is not None else [[] for _ in range(I)]
    if len(instance) != I or any(len(col) > key.u for col in instance):
        return False
    try:
        T = Transcript(proof)
        T.common_scalar(repr_int)
        # a committed column's commitment [p(tau)]G, p its Lagrange polynomial
        inst_com = [g_mul(jac_of_affine(O.generator()),
                          sum(v * _lagrange_at(key, key.tau, i) for i, v in enumerate(col)) % R)
                    for col in instance[:sp.committed]]
        for i, col in enumerate(instance):
            if i < sp.committed:
                T.common_point(inst_com[i])
            else:
                T.common_scalar(len(col))
                for v in col:
                    T.common_scalar(v % R)
        adv_com = [T.read_point() for _ in range(A)]
        theta = T.squeeze()
        lk_perm_com = [(T.read_point(), T.read_point()) for _ in range(L)]
        beta, gamma = T.squeeze(), T.squeeze()
        z_com = [T.read_point() for _ in range(Sn)]
        lz_com = [T.read_point() for _ in range(L)]
        trash = T.squeeze()
        tr_com = [T.read_point() for _ in range(TR)]
        rand_com = T.read_point()
        y = T.squeeze()
        h_com = [T.read_point() for _ in range(key.pieces)]
        x = T.squeeze()
        committed_q = [(col, rot) for col, rot in sp.iq if col < sp.committed]
        inst_ev = [T.read_scalar() for _ in committed_q]
        adv_ev = [T.read_scalar() for _ in sp.aq]
        fixed_ev = [T.read_scalar() for _ in sp.fq]
        rand_ev = T.read_scalar()
        sigma_ev = [T.read_scalar() for _ in range(len(sp.perm))]
        z_ev = []
        for s in range(Sn):
            zx, zn = T.read_scalar(), T.read_scalar()
            zl = T.read_scalar() if s + 1 < Sn else None
            z_ev.append((zx, zn, zl))
        lk_ev = [tuple(T.read_scalar() for _ in range(5)) for _ in range(L)]
        tr_ev = [T.read_scalar() for _ in range(TR)]
    except ValueError:
        return False

    def ev(kind, col, rot):
        if kind == "advice":
            return adv_ev[sp.aq.index((col, rot))]
        if kind == "fixed":
            return fixed_ev[sp.fq.index((col, rot))]
        if col < sp.committed:
            return inst_ev[committed_q.index((col, rot))]
        pt = _rot(key, x, rot)      # instance: interpolated from the values
        return sum(v * _lagrange_at(key, pt, i) for i, v in enumerate(instance[col])) % R

    # numerator at x from the evaluations, folded in y as evaluate_numerator folds it
    l0 = _lagrange_at(key, x, 0)
    l_last = _lagrange_at(key, x, key.u)
    l_blind = sum(_lagrange_at(key, x, i) for i in range(n - bf, n)) % R
    l_active = (1 - l_last - l_blind) % R
    v = 0
    for g in sp.gates:
        v = (v * y + _expr_at(g, ev, key)) % R
    v = (v * y + (1 - z_ev[0][0]) * l0) % R
    zl = z_ev[-1][0]
    v = (v * y + (zl * zl - zl) * l_last) % R
    for s in range(1, Sn):
        v = (v * y + (z_ev[s][0] - z_ev[s - 1][2]) * l0) % R
    col_ev = [ev({0: "fixed", 1: "advice", 2: "instance"}[kind], idx, 0) for kind, idx in sp.perm]
    cur_delta = beta * x % R
    for s in range(Sn):
        left, right = z_ev[s][1], z_ev[s][0]
        for j in range(s * S.CHUNK, min((s + 1) * S.CHUNK, len(sp.perm))):
            left = left * (col_ev[j] + beta * sigma_ev[j] + gamma) % R
            right = right * (col_ev[j] + cur_delta + gamma) % R
            cur_delta = cur_delta * key.delta % R
        v = (v * y + (left - right) * l_active) % R
    for l, (inputs, tables) in enumerate(sp.lookups):       # lookup/verifier.rs
        zx, zn, ai, ai_prev, si = lk_ev[l]
        comp = lambda es: sum(_expr_at(e, ev, key) * pow(theta, len(es) - 1 - j, R) for j, e in enumerate(es)) % R
        v = (v * y + (1 - zx) * l0) % R
        v = (v * y + (zx * zx - zx) * l_last) % R
        lhs = zn * (ai + beta) * (si + gamma)
        rhs = zx * (comp(inputs) + beta) * (comp(tables) + gamma)
        v = (v * y + (lhs - rhs) * l_active) % R
        v = (v * y + (ai - si) * l0) % R
        v = (v * y + (ai - si) * (ai - ai_prev) * l_active) % R
    for t, (sel, cs) in enumerate(sp.trash):                # trash.rs Evaluated::expressions
        comp = 0
        for e in cs:
            comp = (comp * trash + _expr_at(e, ev, key)) % R
        v = (v * y + comp - (1 - ev("fixed", sel, 0)) * tr_ev[t]) % R
    h_ev = v * pow((pow(x, n, R) - 1) % R, R - 2, R) % R

    sf = pow(x, n - 1, R)
    h_comb = O.g1j()
    for i, cm in enumerate(h_com):
        h_comb = g_add(h_comb, g_mul(cm, pow(sf, i, R)))
    x_next, x_last, x_prev = _rot(key, x, 1), _rot(key, x, -(bf + 1)), _rot(key, x, -1)
    queries = [(("adv", col), adv_com[col], _rot(key, x, rot), adv_ev[i]) for i, (col, rot) in enumerate(sp.aq)]
    queries += [(("inst", col), inst_com[col], _rot(key, x, rot), inst_ev[i]) for i, (col, rot) in enumerate(committed_q)]
    for s in range(Sn):
        queries += [(("z", s), z_com[s], x, z_ev[s][0]), (("z", s), z_com[s], x_next, z_ev[s][1])]
    for s in range(Sn - 2, -1, -1):
        queries.append((("z", s), z_com[s], x_last, z_ev[s][2]))
    for l in range(L):
        zx, zn, ai, ai_prev, si = lk_ev[l]
        queries += [(("lz", l), lz_com[l], x, zx), (("li", l), lk_perm_com[l][0], x, ai),
                    (("lt", l), lk_perm_com[l][1], x, si), (("li", l), lk_perm_com[l][0], x_prev, ai_prev),
                    (("lz", l), lz_com[l], x_next, zn)]
    queries += [(("trash", t), tr_com[t], x, tr_ev[t]) for t in range(TR)]
    queries += [(("fixed", col), fixed_com[col], _rot(key, x, rot), fixed_ev[i]) for i, (col, rot) in enumerate(sp.fq)]
    queries += [(("sigma", j), cm, x, sigma_ev[j]) for j, cm in enumerate(sigma_com)]
    queries += [(("h",), h_comb, x, h_ev), (("rand",), rand_com, x, rand_ev)]
    return _multi_open_verify(T, key, queries)


def _multi_open_verify(T, key, queries):
    x1, x2 = T.squeeze(), T.squeeze()
    points, cids, cpoints, cobj, cset, set_points = _intermediate_sets([(q[0], q[1], q[2]) for q in queries])
    evals = {}
    for cid, _, pt, ev in queries:
        evals[(cid, points.index(pt))] = ev
    nsets = len(set_points)
    q_com, q_ev = [], []
    for s in range(nsets):
        members = [c for c in range(len(cids)) if cset[c] == s]
        com = O.g1j()
        evs = [0] * len(set_points[s])
        for j, c in enumerate(members):
            pw = pow(x1, j, R)
            com = g_add(com, g_mul(cobj[c], pw))
            for t, pi in enumerate(set_points[s]):
                evs[t] = (evs[t] + pw * evals[(cids[c], pi)]) % R
        q_com.append(com)
        q_ev.append(evs)
    order = sorted(range(nsets), key=lambda i: (len(set_points[i]), i))
    try:
        f_com = T.read_point()
        x3 = T.squeeze()
        q_x3 = [T.read_scalar() for _ in range(nsets)]
        x4 = T.squeeze()
        pi = T.read_point()
    except ValueError:
        return False
    if T.pos != len(T.rd):
        return False
    f_eval = 0
    for oi in reversed(range(nsets)):
        s = order[oi]
        pts = [points[p] for p in set_points[s]]
        r = 0
        for j, pj in enumerate(pts):
            term = q_ev[s][j]
            for m, pm in enumerate(pts):
                if m != j:
                    term = term * (x3 - pm) * pow((pj - pm) % R, R - 2, R) % R
            r = (r + term) % R
        den = 1
        for pj in pts:
            den = den * (x3 - pj) % R
        f_eval = (f_eval * x2 + (q_x3[oi] - r) * pow(den, R - 2, R)) % R
    final = O.g1j()
    v = 0
    for oi in range(nsets):
        pw = pow(x4, oi, R)
        final = g_add(final, g_mul(q_com[order[oi]], pw))
        v = (v + pw * q_x3[oi]) % R
    pw = pow(x4, nsets, R)
    final = g_add(final, g_mul(f_com, pw))
    v = (v + pw * f_eval) % R
    rhs = g_add(g_add(final, g_neg(g_mul(jac_of_affine(O.generator()), v))), g_mul(pi, x3))
    lhs = g_mul(pi, key.tau)
    return O.g1_eq(lhs, rhs)
