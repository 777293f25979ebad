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
        self.fixed_lag = [m(col) for col in c.cols[5:]]
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
        hv.update(S.DESCRIPTION)
        self.repr = int.from_bytes(hv.digest(), "little") % R


class VerifierKey:
    """What the designated verifier reads of a key, without the prover's
    polynomials: sizes, domain constants and the SRS trapdoor (for the GPU
    keys of circuits too large for CpuKey)."""

    def __init__(self, k, tau_int):
        self.k, self.n = k, 1 << k
        self.tau = tau_int
        self.bf = S.BLINDING
        self.u = self.n - self.bf - 1
        self.pieces = S.DEGREE - 1
        self.omega = S.omega_int(k)
        self.delta = S.fr_const("delta")
        self.nsets = (len(S.PERM_COLUMNS) + S.CHUNK - 1) // S.CHUNK


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


def cpu_prove(key, advice_cols, seed, threads=None, rng_threads=1):
    """advice_cols: list of 5 lists of n ints (Lagrange values); rng_threads:
    the thread count of the reference prover whose random draws to follow."""
    c, n, k, A = key.c, key.n, key.k, S.ADVICE
    Sn = c.nsets
    rng = O.ChaCha(seed=seed)
    if isinstance(advice_cols, np.ndarray):     # Montgomery limbs, column after column
        adv = np.array(advice_cols, dtype=np.uint64, copy=True).reshape(A * n, 4)
    else:
        adv = S.mont([v for col in advice_cols for v in col]).copy()
    for col in range(A):
        adv[col * n + key.u:(col + 1) * n] = rng.random_fr(n - key.u)
    perm_blinds = rng.random_fr(Sn * key.bf)
    rand_poly = sample_random(rng, n, rng_threads)
    limb_blinds = rng.random_fr(key.pieces - 1)

    T = Transcript()
    T.common_scalar(key.repr)
    adv_coeff = [O.ifft(adv[i * n:(i + 1) * n], k, threads) for i in range(A)]
    for p in adv_coeff:
        T.write_point(O.msm(key.srs, p, threads))
    theta, beta, gamma = T.squeeze(), T.squeeze(), T.squeeze()

    cols = np.concatenate([adv] + key.fixed_lag)
    z_lag = O.permutation_products(cols, np.concatenate(key.sigma_lag), n, S.CHUNK, to_fr(beta), to_fr(gamma),
                                   to_fr(key.omega), S.mont(c.delta_pows), perm_blinds, key.bf, threads)
    z_coeff = [O.ifft(z_lag[s * n:(s + 1) * n], k, threads) for s in range(Sn)]
    for p in z_coeff:
        T.write_point(O.msm(key.srs, p, threads))
    trash = T.squeeze()
    T.write_point(O.msm(key.srs, rand_poly, threads))
    y = T.squeeze()

    N = key.N
    adv_ext = np.concatenate([O.coeff_to_extended(p, key.dom, threads) for p in adv_coeff])
    z_ext = np.concatenate([O.coeff_to_extended(p, key.dom, threads) for p in z_coeff])
    nu = O.numerator(N, fixed=key.fixed_ext, advice=adv_ext, beta=to_fr(beta), gamma=to_fr(gamma),
                     theta=to_fr(theta), trash_challenge=to_fr(trash), y=to_fr(y),
                     omega=np.array(key.dom.fr("extended_omega")), g_coset=np.array(key.dom.fr("g_coset")),
                     rot_scale=N // n, last_rotation=-(key.bf + 1), chunk_len=S.CHUNK, l0=key.l0_ext,
                     l_last=key.l_last_ext, l_active=key.l_active_ext, permutation_products=z_ext,
                     permutation_columns=S.PERM_COLUMNS, permutation_pk_cosets=key.sigma_ext,
                     custom_graph=c.custom_graph(), threads=threads)
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

    for col in range(A):
        T.write_scalar(_eval(adv_coeff[col], x))
    for p in key.fixed_coeff:
        T.write_scalar(_eval(p, x))
    T.write_scalar(_eval(rand_poly, x))
    for p in key.sigma_coeff:
        T.write_scalar(_eval(p, x))
    x_next, x_last = _rot(key, x, 1), _rot(key, x, -(key.bf + 1))
    for s in range(Sn):
        T.write_scalar(_eval(z_coeff[s], x))
        T.write_scalar(_eval(z_coeff[s], x_next))
        if s + 1 < Sn:
            T.write_scalar(_eval(z_coeff[s], x_last))
    sf = pow(x, n - 1, R)
    h_poly = _lincomb(h_pieces, [pow(sf, i, R) for i in range(key.pieces)], n)

    queries = [(("adv", i), adv_coeff[i], x) for i in range(A)]
    for s in range(Sn):
        queries += [(("z", s), z_coeff[s], x), (("z", s), z_coeff[s], x_next)]
    for s in range(Sn - 2, -1, -1):
        queries.append((("z", s), z_coeff[s], x_last))
    queries += [(("fixed", i), p, x) for i, p in enumerate(key.fixed_coeff)]
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


def verify(vk, proof, key):
    """vk: (fixed commitments, permutation commitments, transcript repr int).
    Returns True when the proof verifies."""
    fixed_com, sigma_com, repr_int = vk
    n, A, Sn, bf = key.n, S.ADVICE, key.nsets, key.bf
    try:
        T = Transcript(proof)
        T.common_scalar(repr_int)
        adv_com = [T.read_point() for _ in range(A)]
        theta, beta, gamma = T.squeeze(), T.squeeze(), T.squeeze()
        z_com = [T.read_point() for _ in range(Sn)]
        trash = T.squeeze()
        rand_com = T.read_point()
        y = T.squeeze()
        h_com = [T.read_point() for _ in range(key.pieces)]
        x = T.squeeze()
        adv_ev = [T.read_scalar() for _ in range(A)]
        fixed_ev = [T.read_scalar() for _ in range(S.FIXED)]
        rand_ev = T.read_scalar()
        sigma_ev = [T.read_scalar() for _ in range(len(S.PERM_COLUMNS))]
        z_ev = []
        for s in range(Sn):
            zx, zn = T.read_scalar(), T.read_scalar()
            zl = T.read_scalar() if s + 1 < Sn else None
            z_ev.append((zx, zn, zl))
    except ValueError:
        return False

    # numerator at x from the evaluations
    l0 = _lagrange_at(key, x, 0)
    l_last = _lagrange_at(key, x, key.u)
    l_blind = sum(_lagrange_at(key, x, i) for i in range(n - bf, n)) % R
    a, b, c, d = adv_ev[:4]
    q = fixed_ev[0]
    v = q * (a * b + c - d) % R
    v = (v * y + (1 - z_ev[0][0]) * l0) % R
    zl = z_ev[-1][0]
    v = (v * y + (zl * zl - zl) * l_last) % R
    for s in range(1, Sn):
        v = (v * y + (z_ev[s][0] - z_ev[s - 1][2]) * l0) % R
    col_ev = adv_ev + fixed_ev
    cur_delta = beta * x % R
    for s in range(Sn):
        left, right = z_ev[s][1], z_ev[s][0]
        for j in range(s * S.CHUNK, min((s + 1) * S.CHUNK, len(S.PERM_COLUMNS))):
            left = left * (col_ev[j] + beta * sigma_ev[j] + gamma) % R
            right = right * (col_ev[j] + cur_delta + gamma) % R
            cur_delta = cur_delta * key.delta % R
        v = (v * y + (left - right) * (1 - l_last - l_blind)) % R
    h_ev = v * pow((pow(x, n, R) - 1) % R, R - 2, R) % R

    sf = pow(x, n - 1, R)
    h_comb = O.g1j()
    for i, cm in enumerate(h_com):
        h_comb = g_add(h_comb, g_mul(cm, pow(sf, i, R)))
    x_next, x_last = _rot(key, x, 1), _rot(key, x, -(bf + 1))
    queries = [(("adv", i), adv_com[i], x, adv_ev[i]) for i in range(A)]
    for s in range(Sn):
        queries += [(("z", s), z_com[s], x, z_ev[s][0]), (("z", s), z_com[s], x_next, z_ev[s][1])]
    for s in range(Sn - 2, -1, -1):
        queries.append((("z", s), z_com[s], x_last, z_ev[s][2]))
    queries += [(("fixed", i), cm, x, fixed_ev[i]) for i, cm in enumerate(fixed_com)]
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
