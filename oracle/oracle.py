"""ctypes view of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, where it is the checker or the timed CPU
baseline -- never the thing that is measured or shipped.

Arrays are numpy uint64 views of the reference ABI layouts:
Fr (n, 4), G1 affine (n, 12) = x[6] || y[6], G1 Jacobian (n, 18).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
u8p = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class Domain(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("extended_k", ctypes.c_uint32), ("j", ctypes.c_uint32)] + [
        (name, ctypes.c_uint64 * 4)
        for name in ("omega", "omega_inv", "extended_omega", "extended_omega_inv",
                     "ifft_divisor", "extended_ifft_divisor", "g_coset", "g_coset_inv",
                     "barycentric_weight")
    ]

    def fr(self, name):
        return np.array(getattr(self, name)[:], dtype=np.uint64)


class Rng(ctypes.Structure):
    _fields_ = [("key", ctypes.c_uint32 * 8), ("counter", ctypes.c_uint64),
                ("buf", ctypes.c_uint32 * 64), ("idx", ctypes.c_uint32)]


def _declare(L):
    V = None
    sig = {
        "orc_fr_add": (V, [u64p, u64p, u64p]), "orc_fr_sub": (V, [u64p, u64p, u64p]),
        "orc_fr_mul": (V, [u64p, u64p, u64p]), "orc_fr_neg": (V, [u64p, u64p]),
        "orc_fr_inv": (V, [u64p, u64p]), "orc_fr_one": (V, [u64p]),
        "orc_fr_from_u64": (V, [u64p, ctypes.c_uint64]),
        "orc_fr_to_canonical": (V, [u64p, u64p]),
        "orc_fr_from_canonical": (ctypes.c_int, [u64p, u64p]),
        "orc_fr_from_uniform_bytes": (V, [u64p, u8p]),
        "orc_fr_inner_product": (V, [u64p, u64p, u64p, ctypes.c_size_t]),
        "orc_fr_mul_add_many": (V, [u64p, u64p, u64p, u64p, ctypes.c_size_t]),
        "orc_fr_lincomb": (V, [u64p, u64p, u64p, ctypes.c_size_t, ctypes.c_size_t]),
        "orc_pcg32_advance": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64]),
        "orc_pcg32_output": (ctypes.c_uint32, [ctypes.c_uint64]),
        "orc_fr_root_of_unity": (V, [u64p]), "orc_fr_zeta": (V, [u64p]), "orc_fr_delta": (V, [u64p]),
        "orc_fp_add": (V, [u64p, u64p, u64p]), "orc_fp_sub": (V, [u64p, u64p, u64p]),
        "orc_fp_mul": (V, [u64p, u64p, u64p]), "orc_fp_inv": (V, [u64p, u64p]),
        "orc_fp_one": (V, [u64p]),
        "orc_fp_to_canonical": (V, [u64p, u64p]),
        "orc_fp_from_canonical": (ctypes.c_int, [u64p, u64p]),
        "orc_g1_generator": (V, [u64p]),
        "orc_g1j_to_affine": (V, [u64p, u64p]), "orc_g1j_from_affine": (V, [u64p, u64p]),
        "orc_g1j_dbl": (V, [u64p, u64p]), "orc_g1j_add": (V, [u64p, u64p, u64p]),
        "orc_g1j_eq": (ctypes.c_int, [u64p, u64p]),
        "orc_g1a_is_on_curve": (ctypes.c_int, [u64p]),
        "orc_g1_mul": (V, [u64p, u64p, u64p]),
        "orc_g1_batch_to_affine": (V, [u64p, u64p, ctypes.c_size_t]),
        "orc_g1_compress": (V, [u8p, u64p]),
        "orc_g1_decompress": (ctypes.c_int, [u64p, u8p]),
        "orc_g1_decompress_many": (ctypes.c_size_t, [u64p, u8p, ctypes.c_size_t, ctypes.c_int]),
        "orc_g1_from_uncompressed": (ctypes.c_int, [u64p, u8p]),
        "orc_g1_to_uncompressed": (V, [u8p, u64p]),
        "orc_msm_naive": (V, [u64p, u64p, u64p, ctypes.c_size_t]),
        "orc_msm_pippenger": (V, [u64p, u64p, u64p, ctypes.c_size_t, ctypes.c_int]),
        "orc_msm_best": (V, [u64p, u64p, u64p, ctypes.c_size_t, ctypes.c_int]),
        "orc_fp_inv_vartime": (V, [u64p, u64p]),
        "orc_fp_sqrt": (ctypes.c_int, [u64p, u64p]),
        "orc_booth_index": (ctypes.c_int32, [ctypes.c_size_t, ctypes.c_size_t, u8p]),
        "orc_srs_powers": (V, [u64p, u64p, ctypes.c_size_t, ctypes.c_int]),
        "orc_g1_fixed_base_batch": (V, [u64p, u64p, ctypes.c_size_t, ctypes.c_int]),
        "orc_best_fft": (V, [u64p, u64p, ctypes.c_uint32, ctypes.c_int]),
        "orc_omega": (V, [u64p, ctypes.c_uint32]),
        "orc_fft": (V, [u64p, ctypes.c_uint32, ctypes.c_int]),
        "orc_ifft": (V, [u64p, ctypes.c_uint32, ctypes.c_int]),
        "orc_mul_zeta": (V, [u64p, ctypes.c_size_t, u64p, u64p]),
        "orc_mul_repeating": (V, [u64p, ctypes.c_size_t, u64p, ctypes.c_size_t]),
        "orc_mul_scalar": (V, [u64p, ctypes.c_size_t, u64p]),
        "orc_domain_new": (V, [ctypes.POINTER(Domain), u64p, ctypes.c_uint32, ctypes.c_uint32]),
        "orc_coeff_to_extended": (V, [u64p, u64p, ctypes.POINTER(Domain), ctypes.c_int]),
        "orc_extended_to_coeff": (V, [u64p, ctypes.POINTER(Domain), ctypes.c_int]),
        "orc_rng_seed_from_u64": (V, [ctypes.POINTER(Rng), ctypes.c_uint64]),
        "orc_rng_next_u64": (ctypes.c_uint64, [ctypes.POINTER(Rng)]),
        "orc_fr_random_many": (V, [u64p, ctypes.c_size_t, ctypes.POINTER(Rng)]),
        "orc_chacha20_block": (V, [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                   ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32)]),
        "orc_eval_numerator": (V, [u64p, ctypes.c_void_p, ctypes.c_int]),
        "orc_shifted_prefix_product": (V, [u64p, u64p, ctypes.c_size_t, u64p]),
        "orc_fr_sort_canonical": (V, [u64p, u64p, ctypes.c_size_t]),
        "orc_lookup_permute": (ctypes.c_int, [u64p, u64p, u64p, u64p, ctypes.c_size_t]),
        "orc_poseidon_permutation": (V, [u64p, u64p, u64p, ctypes.c_void_p]),
        "orc_poseidon_hash_chain": (V, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, u64p, u64p, u64p,
                                        u64p, u64p, u64p]),
        "orc_poseidon_advice_from_witness": (V, [u64p, u64p, ctypes.c_size_t, ctypes.c_size_t, u64p,
                                                 ctypes.POINTER(ctypes.c_uint16)]),
        "orc_lookup_products": (V, [u64p, u64p, u64p, u64p, u64p, ctypes.c_size_t, ctypes.c_size_t, u64p, u64p,
                                    u64p, ctypes.c_size_t]),
        "orc_poly_eval": (V, [u64p, u64p, ctypes.c_size_t, u64p]),
        "orc_g1j_mul": (V, [u64p, u64p, u64p]),
        "orc_g1j_neg": (V, [u64p, u64p]),
        "orc_kate_division": (V, [u64p, u64p, ctypes.c_size_t, u64p]),
        "orc_rng_from_seed": (V, [ctypes.POINTER(Rng), ctypes.POINTER(ctypes.c_uint8)]),
        "orc_rng_next_u32": (ctypes.c_uint32, [ctypes.POINTER(Rng)]),
        "orc_permutation_products": (V, [u64p, u64p, u64p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                         u64p, u64p, u64p, u64p, u64p, ctypes.c_size_t, ctypes.c_int]),
        "orc_graph_eval_rows": (V, [u64p, ctypes.c_void_p, ctypes.c_void_p, u64p, ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _threads(threads):
    """worker threads of a parallel oracle call: the caller's count, else
    the host's cores up to 16 (a GPU host's share)"""
    return threads or min(16, os.cpu_count() or 1)


def P(a):
    """uint64 pointer to a C-contiguous numpy array"""
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u64p)


def B(a):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u8p)


def fr(n=None):
    return np.zeros((4,) if n is None else (n, 4), dtype=np.uint64)


def g1a(n=None):
    return np.zeros((12,) if n is None else (n, 12), dtype=np.uint64)


def g1j(n=None):
    return np.zeros((18,) if n is None else (n, 18), dtype=np.uint64)


# ---- convenience wrappers -------------------------------------------------

def fr_from_int(v):
    limbs = np.array([(v >> (64 * i)) & (2**64 - 1) for i in range(4)], dtype=np.uint64)
    out = fr()
    if not lib().orc_fr_from_canonical(P(out), P(limbs)):
        raise ValueError("not reduced")
    return out


def fr_to_int(a):
    out = fr()
    lib().orc_fr_to_canonical(P(out), P(np.ascontiguousarray(a, dtype=np.uint64)))
    return sum(int(out[i]) << (64 * i) for i in range(4))


def fr_op(name, *args):
    out = fr()
    getattr(lib(), "orc_fr_" + name)(P(out), *[P(np.ascontiguousarray(a)) for a in args])
    return out


def fp_op(name, *args):
    out = np.zeros(6, dtype=np.uint64)
    getattr(lib(), "orc_fp_" + name)(P(out), *[P(np.ascontiguousarray(a)) for a in args])
    return out


def lincomb(polys, scalars, n):
    """sum_j scalars[j] * polys[j], coefficient by coefficient (m x n values)"""
    p = np.ascontiguousarray(polys, dtype=np.uint64).reshape(-1, 4)
    sc = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    assert p.shape[0] == sc.shape[0] * n
    out = fr(n)
    lib().orc_fr_lincomb(P(out), P(p), P(sc), sc.shape[0], n)
    return out


def mul_add(a, b, c):
    """a * b + c element by element (Montgomery arrays of one length)"""
    a, b, c = (np.ascontiguousarray(x, dtype=np.uint64) for x in (a, b, c))
    out = fr(a.shape[0])
    lib().orc_fr_mul_add_many(P(out), P(a), P(b), P(c), a.shape[0])
    return out


def inner_product(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = fr()
    lib().orc_fr_inner_product(P(out), P(a), P(b), a.shape[0])
    return out


def g1_eq(p, q):
    """equality of two Jacobian points (any Z)"""
    return bool(lib().orc_g1j_eq(P(np.ascontiguousarray(p, dtype=np.uint64)),
                                 P(np.ascontiguousarray(q, dtype=np.uint64))))


def decompress(buf):
    b = np.frombuffer(bytes(buf), dtype=np.uint8).copy()
    out = g1a()
    if not lib().orc_g1_decompress(P(out), B(b)):
        raise ValueError("invalid compressed point")
    return out


def generator():
    g = g1a()
    lib().orc_g1_generator(P(g))
    return g


def to_affine(j):
    j = np.ascontiguousarray(j, dtype=np.uint64)
    if j.ndim == 1:
        out = g1a()
        lib().orc_g1j_to_affine(P(out), P(j))
        return out
    out = g1a(j.shape[0])
    lib().orc_g1_batch_to_affine(P(out), P(j), j.shape[0])
    return out


def g1_mul(p_aff, s):
    out = g1j()
    lib().orc_g1_mul(P(out), P(np.ascontiguousarray(p_aff)), P(np.ascontiguousarray(s)))
    return out


def random_fr(n, seed):
    rng = Rng()
    lib().orc_rng_seed_from_u64(ctypes.byref(rng), seed)
    out = fr(n)
    if n:
        lib().orc_fr_random_many(P(out), n, ctypes.byref(rng))
    return out


def msm(points, scalars, threads=None, naive=False, best=False):
    """naive: double-and-add per point; best: msm_best with its batch-affine
    schedule; else Pippenger with Jacobian buckets"""
    points = np.ascontiguousarray(points, dtype=np.uint64)
    scalars = np.ascontiguousarray(scalars, dtype=np.uint64)
    n = points.shape[0]
    assert scalars.shape[0] == n
    out = g1j()
    if naive:
        lib().orc_msm_naive(P(out), P(points), P(scalars), n)
    elif best:
        lib().orc_msm_best(P(out), P(points), P(scalars), n, _threads(threads))
    else:
        lib().orc_msm_pippenger(P(out), P(points), P(scalars), n, _threads(threads))
    return out


def fixed_base(ts, threads=None):
    ts = np.ascontiguousarray(ts, dtype=np.uint64)
    out = g1a(ts.shape[0])
    lib().orc_g1_fixed_base_batch(P(out), P(ts), ts.shape[0], _threads(threads))
    return out


def srs(tau, n, threads=None):
    out = g1a(n)
    lib().orc_srs_powers(P(out), P(np.ascontiguousarray(tau)), n, _threads(threads))
    return out


def omega(log_n):
    out = fr()
    lib().orc_omega(P(out), log_n)
    return out


def fft(a, log_n, threads=None):
    a = np.array(a, dtype=np.uint64, order="C", copy=True)
    lib().orc_fft(P(a), log_n, _threads(threads))
    return a


def ifft(a, log_n, threads=None):
    a = np.array(a, dtype=np.uint64, order="C", copy=True)
    lib().orc_ifft(P(a), log_n, _threads(threads))
    return a


def best_fft(a, w, log_n, threads=None):
    a = np.array(a, dtype=np.uint64, order="C", copy=True)
    lib().orc_best_fft(P(a), P(np.ascontiguousarray(w)), log_n, _threads(threads))
    return a


def domain(j, k):
    d = Domain()
    ek = k
    while (1 << ek) < (1 << k) * (j - 1):
        ek += 1
    t = fr(1 << (ek - k))
    lib().orc_domain_new(ctypes.byref(d), P(t), j, k)
    return d, t


def mul_zeta(a, first, second):
    a = np.array(a, dtype=np.uint64, order="C", copy=True)
    lib().orc_mul_zeta(P(a), a.shape[0], P(np.ascontiguousarray(first)), P(np.ascontiguousarray(second)))
    return a


def mul_repeating(a, factors):
    a = np.array(a, dtype=np.uint64, order="C", copy=True)
    f = np.ascontiguousarray(factors, dtype=np.uint64)
    lib().orc_mul_repeating(P(a), a.shape[0], P(f), f.shape[0])
    return a


def mul_scalar(a, s):
    a = np.array(a, dtype=np.uint64, order="C", copy=True)
    lib().orc_mul_scalar(P(a), a.shape[0], P(np.ascontiguousarray(s)))
    return a


def coeff_to_extended(coeffs, d, threads=None):
    n = 1 << d.k
    N = 1 << d.extended_k
    assert coeffs.shape[0] == n
    out = fr(N)
    lib().orc_coeff_to_extended(P(out), P(np.ascontiguousarray(coeffs)), ctypes.byref(d), _threads(threads))
    return out


def extended_to_coeff(values, d, threads=None):
    a = np.array(values, dtype=np.uint64, order="C", copy=True)
    lib().orc_extended_to_coeff(P(a), ctypes.byref(d), _threads(threads))
    return a


def compress(p_aff):
    out = np.zeros(48, dtype=np.uint8)
    lib().orc_g1_compress(B(out), P(np.ascontiguousarray(p_aff)))
    return bytes(out)


def to_uncompressed(p_aff):
    out = np.zeros(96, dtype=np.uint8)
    lib().orc_g1_to_uncompressed(B(out), P(np.ascontiguousarray(p_aff)))
    return bytes(out)


def fp_sqrt(a):
    """(root, is_square) of a Montgomery Fp element"""
    out = np.zeros(6, dtype=np.uint64)
    ok = lib().orc_fp_sqrt(P(out), P(np.ascontiguousarray(a, dtype=np.uint64)))
    return out, bool(ok)


def from_uncompressed(buf):
    b = np.frombuffer(bytes(buf), dtype=np.uint8).copy()
    out = g1a()
    ok = lib().orc_g1_from_uncompressed(P(out), B(b))
    return out if ok else None


def load_srs_file(path):
    """Parse a bls_midnight_2pK file (params.rs:170-262, RawBytes format):
    u32 k || 2^k uncompressed G1 g || 2^k uncompressed G1 g_lagrange || G2 || sG2.
    Returns (k, g, g_lagrange) as affine Montgomery arrays."""
    with open(path, "rb") as f:
        data = f.read()
    k = int.from_bytes(data[:4], "little")
    n = 1 << k
    body = np.frombuffer(data, dtype=np.uint8, count=2 * n * 96, offset=4).reshape(2 * n, 96)
    pts = g1a(2 * n)
    for i in range(2 * n):
        row = np.ascontiguousarray(body[i])
        if not lib().orc_g1_from_uncompressed(P(pts[i]), B(row)):
            raise ValueError("invalid point %d" % i)
    return k, pts[:n].copy(), pts[n:].copy()


def load_srs_fixture(path):
    """Read tests/golden/srs_2p14.g1c: u32 k || 2^k compressed g || 2^k compressed g_lagrange."""
    with open(path, "rb") as f:
        data = f.read()
    k = int.from_bytes(data[:4], "little")
    n = 1 << k
    raw = np.frombuffer(data, dtype=np.uint8, offset=4, count=2 * n * 48).copy()
    pts = g1a(2 * n)
    bad = lib().orc_g1_decompress_many(P(pts), B(raw), 2 * n, os.cpu_count())
    if bad:
        raise ValueError("%d invalid points in %s" % (bad, path))
    return k, pts[:n].copy(), pts[n:].copy()


# ---------------------------------------------------------------- quotient numerator

class OrcValueSource(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("a", ctypes.c_uint32), ("b", ctypes.c_uint32)]


class OrcGraph(ctypes.Structure):
    _fields_ = [("constants", ctypes.c_void_p), ("constants_len", ctypes.c_size_t),
                ("rotations", ctypes.c_void_p), ("rotations_len", ctypes.c_size_t),
                ("calculations", ctypes.c_void_p), ("calculations_len", ctypes.c_size_t),
                ("horner_parts", ctypes.c_void_p), ("horner_parts_len", ctypes.c_size_t),
                ("num_intermediates", ctypes.c_uint32)]


class OrcColumnRef(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("index", ctypes.c_uint32)]


_FR4 = ctypes.c_uint64 * 4


class NumeratorInputs(ctypes.Structure):
    _fields_ = [
        ("len", ctypes.c_size_t), ("run_custom", ctypes.c_int), ("initial_out", ctypes.c_void_p),
        ("fixed", ctypes.c_void_p), ("fixed_cols", ctypes.c_size_t),
        ("advice", ctypes.c_void_p), ("advice_cols", ctypes.c_size_t),
        ("instance", ctypes.c_void_p), ("instance_cols", ctypes.c_size_t),
        ("challenges", ctypes.c_void_p), ("challenge_len", ctypes.c_size_t),
        ("beta", _FR4), ("gamma", _FR4), ("theta", _FR4), ("trash_challenge", _FR4), ("y", _FR4),
        ("omega", _FR4), ("g_coset", _FR4),
        ("rot_scale", ctypes.c_int32), ("last_rotation", ctypes.c_int32), ("chunk_len", ctypes.c_uint32),
        ("l0", ctypes.c_void_p), ("l_last", ctypes.c_void_p), ("l_active", ctypes.c_void_p),
        ("permutation_products", ctypes.c_void_p), ("permutation_sets", ctypes.c_size_t),
        ("permutation_columns", ctypes.c_void_p), ("permutation_columns_len", ctypes.c_size_t),
        ("permutation_pk_cosets", ctypes.c_void_p),
        ("lookup_product_cosets", ctypes.c_void_p), ("lookup_input_cosets", ctypes.c_void_p),
        ("lookup_table_cosets", ctypes.c_void_p), ("lookup_count", ctypes.c_size_t),
        ("trash_cosets", ctypes.c_void_p), ("trash_selector_fixed_indices", ctypes.c_void_p),
        ("trash_count", ctypes.c_size_t),
        ("custom_graph", OrcGraph), ("lookup_graphs", ctypes.c_void_p), ("trash_graphs", ctypes.c_void_p),
    ]


def _orc_graph(g):
    """Any graph object with as_raw() (RawGraph layout) -> OrcGraph."""
    if g is None:
        return OrcGraph()
    r = g.as_raw()
    return OrcGraph(r.constants, r.constants_len, r.rotations, r.rotations_len, r.calculations,
                    r.calculations_len, r.horner_parts, r.horner_parts_len, r.num_intermediates)


def _col(a, n):
    if a is None:
        return None, 0
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4)
    assert a.shape[0] % n == 0
    return a, a.shape[0] // n


def _set4(field, v):
    v = np.asarray(v, dtype=np.uint64).reshape(4)
    for i in range(4):
        field[i] = int(v[i])


def numerator(n, *, beta, gamma, theta, trash_challenge, y, omega, g_coset, rot_scale, last_rotation,
              chunk_len, custom_graph=None, run_custom=True, initial_out=None, fixed=None, advice=None,
              instance=None, challenges=None, l0=None, l_last=None, l_active=None, permutation_products=None,
              permutation_columns=(), permutation_pk_cosets=None, lookup_product_cosets=None,
              lookup_input_cosets=None, lookup_table_cosets=None, lookup_graphs=(), trash_cosets=None,
              trash_selector_fixed_indices=(), trash_graphs=(), threads=None):
    """orc_eval_numerator: evaluate_numerator for one proof, over n rows.
    g_coset multiplies beta into the first identity column's delta."""
    keep = []

    def ptr(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data

    s = NumeratorInputs()
    s.len = n
    s.run_custom = 1 if run_custom else 0
    s.initial_out = ptr(None if initial_out is None else np.asarray(initial_out, dtype=np.uint64))
    for name, arr in (("fixed", fixed), ("advice", advice), ("instance", instance)):
        a, c = _col(arr, n)
        setattr(s, name, ptr(a))
        setattr(s, name + "_cols", c)
    if challenges is not None:
        ch = np.ascontiguousarray(challenges, dtype=np.uint64).reshape(-1, 4)
        s.challenges = ptr(ch)
        s.challenge_len = ch.shape[0]
    for name, v in (("beta", beta), ("gamma", gamma), ("theta", theta), ("trash_challenge", trash_challenge),
                    ("y", y), ("omega", omega), ("g_coset", g_coset)):
        _set4(getattr(s, name), v)
    s.rot_scale, s.last_rotation, s.chunk_len = rot_scale, last_rotation, chunk_len
    for name, v in (("l0", l0), ("l_last", l_last), ("l_active", l_active)):
        setattr(s, name, ptr(None if v is None else np.asarray(v, dtype=np.uint64)))
    pp, nsets = _col(permutation_products, n)
    s.permutation_products, s.permutation_sets = ptr(pp), nsets
    refs = list(permutation_columns)
    if refs:
        arr = (OrcColumnRef * len(refs))(*[OrcColumnRef(k, i) for k, i in refs])
        keep.append(arr)
        s.permutation_columns = ctypes.addressof(arr)
    s.permutation_columns_len = len(refs)
    s.permutation_pk_cosets = ptr(_col(permutation_pk_cosets, n)[0])
    lp, nlook = _col(lookup_product_cosets, n)
    s.lookup_product_cosets, s.lookup_count = ptr(lp), nlook
    s.lookup_input_cosets = ptr(_col(lookup_input_cosets, n)[0])
    s.lookup_table_cosets = ptr(_col(lookup_table_cosets, n)[0])
    tc, ntrash = _col(trash_cosets, n)
    s.trash_cosets, s.trash_count = ptr(tc), ntrash
    if ntrash:
        s.trash_selector_fixed_indices = ptr(np.asarray(trash_selector_fixed_indices, dtype=np.uint32))
    s.custom_graph = _orc_graph(custom_graph)
    for name, gs in (("lookup_graphs", lookup_graphs), ("trash_graphs", trash_graphs)):
        gs = list(gs)
        if gs:
            arr = (OrcGraph * len(gs))(*[_orc_graph(g) for g in gs])
            keep.append(arr)
            setattr(s, name, ctypes.addressof(arr))
    out = np.zeros((n, 4), dtype=np.uint64)
    lib().orc_eval_numerator(P(out), ctypes.addressof(s), _threads(threads))
    return out


def graph_rows(n, graph, *, beta=None, gamma=None, theta=None, trash_challenge=None, y=None, rot_scale=1,
               fixed=None, advice=None, instance=None, challenges=None, previous=None, threads=None):
    """orc_graph_eval_rows: one graph over n rows (previous: per-row
    PreviousValue or None for zero)."""
    z = np.zeros(4, dtype=np.uint64)
    kw = dict(beta=z if beta is None else beta, gamma=z if gamma is None else gamma,
              theta=z if theta is None else theta,
              trash_challenge=z if trash_challenge is None else trash_challenge, y=z if y is None else y)
    keep = []
    s = NumeratorInputs()
    s.len = n
    for name, arr in (("fixed", fixed), ("advice", advice), ("instance", instance)):
        a, c = _col(arr, n)
        keep.append(a)
        setattr(s, name, None if a is None else a.ctypes.data)
        setattr(s, name + "_cols", c)
    if challenges is not None:
        ch = np.ascontiguousarray(challenges, dtype=np.uint64).reshape(-1, 4)
        keep.append(ch)
        s.challenges, s.challenge_len = ch.ctypes.data, ch.shape[0]
    for name, v in kw.items():
        _set4(getattr(s, name), v)
    s.rot_scale = rot_scale
    g = _orc_graph(graph)
    prev = None if previous is None else np.ascontiguousarray(previous, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros((n, 4), dtype=np.uint64)
    lib().orc_graph_eval_rows(P(out), ctypes.addressof(g), ctypes.addressof(s),
                              None if prev is None else P(prev), _threads(threads))
    return out


def permutation_products(columns, sigmas, n, chunk_len, beta, gamma, omega, delta_powers, blinds, blinding,
                         threads=None):
    """orc_permutation_products: (sets * n, 4) grand products."""
    cols = np.ascontiguousarray(columns, dtype=np.uint64).reshape(-1, 4)
    sig = np.ascontiguousarray(sigmas, dtype=np.uint64).reshape(-1, 4)
    ncol = cols.shape[0] // n
    sets = -(-ncol // chunk_len)
    dp = np.ascontiguousarray(delta_powers, dtype=np.uint64).reshape(-1, 4)
    bl = np.ascontiguousarray(blinds, dtype=np.uint64).reshape(-1, 4) if blinding else fr(1)
    out = fr(sets * n)
    lib().orc_permutation_products(P(out), P(cols), P(sig), n, ncol, chunk_len,
                                   P(np.ascontiguousarray(beta, dtype=np.uint64)),
                                   P(np.ascontiguousarray(gamma, dtype=np.uint64)),
                                   P(np.ascontiguousarray(omega, dtype=np.uint64)), P(dp), P(bl), blinding,
                                   _threads(threads))
    return out


def sort_canonical(values):
    v = np.ascontiguousarray(values, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros_like(v)
    lib().orc_fr_sort_canonical(P(out), P(v), v.shape[0])
    return out


def lookup_permute(inputs, table, usable):
    """(permuted input, permuted table) over the first `usable` rows, or None if an input is not in the table"""
    a = np.ascontiguousarray(inputs, dtype=np.uint64).reshape(-1, 4)
    t = np.ascontiguousarray(table, dtype=np.uint64).reshape(-1, 4)
    pin = np.zeros((usable, 4), dtype=np.uint64)
    ptab = np.zeros((usable, 4), dtype=np.uint64)
    if not lib().orc_lookup_permute(P(pin), P(ptab), P(a), P(t), usable):
        return None
    return pin, ptab


def _fr_arr(x, count=None):
    a = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 4)
    if count is not None and a.shape[0] != count:
        raise ValueError(f"expected {count} field elements, got {a.shape[0]}")
    return a


def poseidon_permutation(state, mds, rc, witness=False):
    """orc_poseidon_permutation: the permuted (3, 4) state, and its 128 witness values when asked"""
    st = np.array(_fr_arr(state, 3), copy=True)
    w = np.zeros((128, 4), dtype=np.uint64) if witness else None
    lib().orc_poseidon_permutation(P(st), P(_fr_arr(mds, 9)), P(_fr_arr(rc, 204)),
                                   None if w is None else w.ctypes.data)
    return (st, w) if witness else st


def poseidon_hash_chain(length, initial, tail0, tail1, input_len, mds, rc, witness=False):
    """orc_poseidon_hash_chain: (length, 4) outputs, or (length * 256, 4) witness values"""
    out = np.zeros((length, 4), dtype=np.uint64)
    w = np.zeros((length * 256, 4), dtype=np.uint64) if witness else None
    args = [np.ascontiguousarray(v, dtype=np.uint64) for v in (initial, tail0, tail1, input_len)]
    lib().orc_poseidon_hash_chain(out.ctypes.data, None if w is None else w.ctypes.data, length,
                                  *[P(a) for a in args], P(_fr_arr(mds, 9)), P(_fr_arr(rc, 204)))
    return w if witness else out


def poseidon_advice_from_witness(witness, hash_count, domain_len, metadata, layout):
    """orc_poseidon_advice_from_witness: (8 * domain_len, 4) column-major advice values"""
    w = _fr_arr(witness, hash_count * 256)
    lay = np.ascontiguousarray(layout, dtype=np.uint16)
    assert lay.shape == (128,)
    out = np.zeros((8 * domain_len, 4), dtype=np.uint64)
    lib().orc_poseidon_advice_from_witness(P(out), P(w), hash_count, domain_len, P(_fr_arr(metadata, 11)),
                                           lay.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    return out


def lookup_products(ci, ct, pi, pt, n, count, beta, gamma, blinds, blinding):
    """orc_lookup_products: count * n values (lookup-major)"""
    arrs = [np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 4) for x in (ci, ct, pi, pt)]
    bl = np.ascontiguousarray(blinds, dtype=np.uint64).reshape(-1, 4) if blinding else np.zeros((1, 4), np.uint64)
    out = np.zeros((n * count, 4), dtype=np.uint64)
    lib().orc_lookup_products(P(out), *[P(a) for a in arrs], n, count, P(np.ascontiguousarray(beta)),
                              P(np.ascontiguousarray(gamma)), P(bl), blinding)
    return out


def shifted_prefix_product(values, initial):
    v = np.ascontiguousarray(values, dtype=np.uint64).reshape(-1, 4)
    out = fr(v.shape[0])
    lib().orc_shifted_prefix_product(P(out), P(v), v.shape[0], P(np.ascontiguousarray(initial, dtype=np.uint64)))
    return out


def poly_eval(coeffs, point):
    c = np.ascontiguousarray(coeffs, dtype=np.uint64).reshape(-1, 4)
    out = fr()
    lib().orc_poly_eval(P(out), P(c), c.shape[0], P(np.ascontiguousarray(point, dtype=np.uint64)))
    return out


def kate_division(coeffs, point):
    """(a - a(b)) / (X - b) with the same length (top coefficient zero)"""
    c = np.array(coeffs, dtype=np.uint64, order="C", copy=True).reshape(-1, 4)
    lib().orc_kate_division(P(c), P(c), c.shape[0], P(np.ascontiguousarray(point, dtype=np.uint64)))
    return c


class ChaCha:
    """The oracle's rand_chacha ChaCha20Rng (seed_from_u64 or a 32-byte seed)."""

    def __init__(self, seed=None, seed_bytes=None):
        self.r = Rng()
        if seed_bytes is not None:
            b = (ctypes.c_uint8 * 32)(*seed_bytes)
            lib().orc_rng_from_seed(ctypes.byref(self.r), b)
        else:
            lib().orc_rng_seed_from_u64(ctypes.byref(self.r), seed)

    def random_fr(self, n):
        out = fr(n)
        if n:
            lib().orc_fr_random_many(P(out), n, ctypes.byref(self.r))
        return out

    def fill_bytes(self, n):
        out = bytearray()
        while len(out) < n:
            w = lib().orc_rng_next_u32(ctypes.byref(self.r))
            out += int(w).to_bytes(4, "little")
        return bytes(out[:n])
