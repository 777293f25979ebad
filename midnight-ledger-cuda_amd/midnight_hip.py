"""ctypes binding of libmidnight_hip.so (the gfx950 backend's C ABI).

Mirrors the Rust-side safe wrappers of midnight-sppark-msm/src/lib.rs
(bls12_381_fr_ntt, msm contexts, device vectors): a nonzero RustError code
raises MidnightHipError carrying the backend's message, as the Rust side
maps it to MsmError::Sppark (lib.rs:86-115).  There is no CPU fallback: if
the library is missing this module raises at load time.

Arrays are numpy uint64 views of the ABI layouts: RawFr (n, 4),
RawG1Affine (n, 12), RawG1Projective (n, 18).
"""
import ctypes
import os

import numpy as np

from numerator_graph import RawColumnRef, RawGraph, empty_raw_graph

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmidnight_hip.so")

u64p = ctypes.POINTER(ctypes.c_uint64)
c_size = ctypes.c_size_t


class MidnightHipError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("midnight_hip error %d: %s" % (code, message))
        self.code = code
        self.message = message


class RustError(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int), ("message", ctypes.c_void_p)]


class RawFr(ctypes.Structure):
    _fields_ = [("limbs", ctypes.c_uint64 * 4)]

    @classmethod
    def of(cls, arr):
        r = cls()
        a = np.ascontiguousarray(arr, dtype=np.uint64).reshape(4)
        for i in range(4):
            r.limbs[i] = int(a[i])
        return r


class RawMsmConfig(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "domain_k", "min_msm_scalars", "unfiltered_min_scalars", "window_bits",
        "sort_grid_divisor", "sort_window_lanes", "integration_blocks", "batch_lanes",
        "reuse_event", "device_reduce")]


class RawMsmSlice(ctypes.Structure):
    _fields_ = [("scalar_offset", c_size), ("point_offset", c_size), ("len", c_size),
                ("nonzero_len", c_size)]


_lib = None
_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MidnightHipError(-1, "libmidnight_hip.so is not built (%s)" % LIB_PATH)
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    E = RustError
    vp = ctypes.c_void_p
    vpp = ctypes.POINTER(ctypes.c_void_p)
    b = ctypes.c_bool
    u32 = ctypes.c_uint32
    sig = {
        "midnight_sppark_execution_lane_set": (E, [u32]),
        "midnight_sppark_bls12_381_g1_msm": (E, [u64p, u64p, u64p, c_size]),
        "midnight_sppark_bls12_381_g1_msm_batch": (E, [u64p, u64p, u64p, c_size, c_size]),
        "midnight_sppark_bls12_381_g1_msm_context_create_with_config": (E, [vpp, u64p, c_size, RawMsmConfig]),
        "midnight_sppark_bls12_381_g1_msm_context_create": (E, [vpp, u64p, c_size]),
        "midnight_sppark_bls12_381_g1_msm_context_free": (None, [vp]),
        "midnight_sppark_bls12_381_g1_msm_context_batch": (E, [u64p, vp, u64p, c_size, c_size]),
        "midnight_sppark_bls12_381_g1_msm_context_device_batch": (E, [u64p, vp, vp, c_size, c_size, c_size]),
        "midnight_sppark_bls12_381_g1_msm_context_device_ragged": (
            E, [u64p, vp, vp, ctypes.POINTER(RawMsmSlice), c_size]),
        "midnight_sppark_bls12_381_g1_msm_context_device_scrambled_batch": (
            E, [u64p, vp, vp, c_size, c_size, c_size, vp]),
        "midnight_sppark_bls12_381_fr_ntt": (E, [u64p, u32, b]),
        "midnight_sppark_bls12_381_fr_ntt_zeta": (E, [u64p, u32, b, RawFr, RawFr, RawFr]),
        "midnight_sppark_bls12_381_fr_divide_ntt_zeta": (E, [u64p, u32, u64p, c_size, RawFr, RawFr, RawFr]),
        "midnight_sppark_bls12_381_fr_coeff_to_extended_batch": (
            E, [u64p, u64p, c_size, c_size, u32, RawFr, RawFr]),
        "midnight_sppark_bls12_381_fr_device_coeff_to_extended_batch": (
            E, [vpp, u64p, c_size, c_size, u32, RawFr, RawFr]),
        "midnight_sppark_bls12_381_fr_device_coeff_to_extended_from_device_batch": (
            E, [vpp, vp, c_size, c_size, u32, RawFr, RawFr]),
        "midnight_sppark_bls12_381_fr_device_vec_ntt_zeta": (E, [vp, u32, b, RawFr, RawFr, RawFr]),
        "midnight_sppark_bls12_381_fr_device_vec_ntt_zeta_batch": (E, [vp, u32, c_size, b, RawFr, RawFr, RawFr]),
        "midnight_sppark_bls12_381_fr_device_vec_divide_ntt_zeta": (
            E, [vp, u32, u64p, c_size, RawFr, RawFr, RawFr]),
        "midnight_sppark_bls12_381_fr_mul_assign": (E, [u64p, RawFr, c_size]),
        "midnight_sppark_bls12_381_fr_mul_assign_zeta": (E, [u64p, RawFr, RawFr, c_size]),
        "midnight_sppark_bls12_381_fr_mul_assign_repeating": (E, [u64p, u64p, c_size, c_size]),
        "midnight_sppark_bls12_381_fr_device_vec_upload": (E, [vpp, u64p, c_size]),
        "midnight_sppark_bls12_381_fr_device_vec_upload_many": (
            E, [vpp, ctypes.POINTER(u64p), ctypes.POINTER(c_size), c_size]),
        "midnight_sppark_bls12_381_fr_device_vec_download": (E, [u64p, vp, c_size]),
        "midnight_sppark_bls12_381_fr_device_vec_free": (None, [vp]),
        "midnight_sppark_bls12_381_fr_device_quotient_pieces": (E, [vpp, vp, c_size, c_size, u64p, c_size]),
        "midnight_hip_fr_device_vec_len": (c_size, [vp]),
        "midnight_hip_g1_msm_context_device_multi": (
            E, [u64p, vp, vp, ctypes.POINTER(c_size), ctypes.POINTER(c_size), c_size]),
        "midnight_hip_fr_device_vec_ntt_multi": (
            E, [vp, ctypes.POINTER(c_size), ctypes.POINTER(u32), c_size, b]),
        "midnight_hip_valu_peak": (
            E, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
        "midnight_hip_copy_bandwidth": (E, [c_size, ctypes.POINTER(ctypes.c_double)]),
        "midnight_hip_fr29_product_latency": (E, [ctypes.POINTER(ctypes.c_double)]),
        "midnight_hip_batch_plan_create": (
            E, [vpp, vp, vp, ctypes.POINTER(c_size), ctypes.POINTER(u32), c_size, u32]),
        "midnight_hip_batch_plan_run": (E, [vp, u64p]),
        "midnight_hip_batch_plan_info": (
            E, [vp, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32),
                ctypes.POINTER(ctypes.c_float)]),
        "midnight_hip_batch_plan_free": (None, [vp]),
        "midnight_hip_field_mul_rate": (
            E, [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "midnight_hip_fr_device_vec_to_montgomery": (E, [vp]),
        "midnight_hip_g1_msm_context_create_from_multiples": (E, [vpp, vp]),
        "midnight_hip_g1_msm_context_points_download": (E, [u64p, vp, c_size]),
        "midnight_hip_g1_msm_context_last_stats": (
            E, [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32)]),
        "midnight_sppark_bls12_381_eval_numerator": (
            E, [u64p, c_size, vp, u32, vp, c_size, vp, c_size, vp, c_size, vp, c_size]
            + [RawFr] * 8 + [ctypes.c_int32, ctypes.c_int32, u32, vp, vp, vp, vp, c_size, vp, c_size, vp, c_size,
                             vp, vp, vp, c_size, vp, vp, c_size, RawGraph, vp, vp]),
        "midnight_sppark_bls12_381_eval_graph_batch": (
            E, [u64p, c_size, vp, c_size, vp, c_size, vp, c_size, vp, c_size] + [RawFr] * 5
            + [ctypes.c_int32, vp, c_size]),
        "midnight_sppark_bls12_381_numerator_context_create": (
            E, [vpp, c_size, c_size, vp, vp, c_size, vp, vp, vp, vp, c_size, vp, vp, c_size, vp, c_size,
                RawGraph, vp, vp, vp, c_size, vp, RawFr, RawFr, RawFr, RawFr, ctypes.c_int32, ctypes.c_int32, u32]),
        "midnight_sppark_bls12_381_numerator_context_free": (None, [vp]),
        "midnight_sppark_bls12_381_numerator_context_eval": (
            E, [vpp, vp, vp, c_size, c_size, c_size, c_size, c_size, vp, c_size] + [RawFr] * 5
            + [c_size] * 8),
        "midnight_sppark_bls12_381_fr_shifted_prefix_product": (E, [u64p, RawFr, c_size]),
        "midnight_sppark_bls12_381_fr_permutation_products": (
            E, [vpp, u64p, u64p, c_size, c_size, c_size, RawFr, RawFr, RawFr, u64p, vp, c_size]),
        "midnight_sppark_bls12_381_fr_device_vec_scatter_host": (E, [vp, ctypes.POINTER(u32), vp, c_size]),
        "midnight_sppark_bls12_381_fr_device_vec_concat_host": (E, [vpp, vp, vp, c_size]),
        "midnight_sppark_bls12_381_fr_device_vec_concat_many_host": (E, [vpp, vpp, c_size, vp, c_size]),
        "midnight_sppark_bls12_381_fr_device_vec_gather_polys": (E, [vpp, vp, c_size, ctypes.POINTER(u32), c_size]),
        "midnight_sppark_bls12_381_fr_device_vec_download_nonzero_suffixes": (
            E, [u64p, c_size, ctypes.POINTER(c_size), ctypes.POINTER(c_size), ctypes.POINTER(c_size), vp, c_size,
                c_size, c_size]),
        "midnight_sppark_bls12_381_fr_poly_lincomb": (E, [u64p, u64p, c_size, c_size, u64p]),
        "midnight_sppark_bls12_381_fr_device_poly_lincomb": (E, [vpp, vp, c_size, c_size, u64p, c_size]),
        "midnight_sppark_bls12_381_fr_device_eval_polys": (E, [u64p, vp, c_size, c_size, RawFr]),
        "midnight_sppark_bls12_381_fr_multiopen_prepare": (
            E, [u64p, u64p, u64p, c_size, c_size, ctypes.POINTER(u32), c_size, ctypes.POINTER(u32), u64p, c_size,
                u64p, c_size, ctypes.POINTER(u32), ctypes.POINTER(u32), c_size]),
        "midnight_sppark_bls12_381_fr_multiopen_prepare_device": (
            E, [vp, vpp, u64p, vp, vp, vp, c_size, ctypes.POINTER(RawDevicePolyRef), c_size, ctypes.POINTER(u32),
                c_size, ctypes.POINTER(u32), u64p, c_size, u64p, c_size, ctypes.POINTER(u32), ctypes.POINTER(u32),
                c_size]),
        "midnight_sppark_bls12_381_fr_multiopen_f_poly_device": (
            E, [vpp, vp, c_size, ctypes.POINTER(u32), c_size, u64p, c_size, u64p, c_size]),
        "midnight_sppark_bls12_381_fr_multiopen_pi_device": (E, [vpp, vp, vp, c_size, u64p, c_size, RawFr]),
        "midnight_sppark_bls12_381_fr_invert_assign": (E, [u64p, c_size]),
        "midnight_sppark_bls12_381_fr_lookup_products": (
            E, [vpp, u64p, u64p, u64p, u64p, c_size, c_size, RawFr, RawFr, vp, c_size]),
        "midnight_sppark_bls12_381_fr_lookup_products_device": (
            E, [vpp, vp, c_size, c_size, RawFr, RawFr, vp, c_size]),
        "midnight_sppark_bls12_381_fr_sort_canonical": (E, [u64p, c_size]),
        "midnight_sppark_bls12_381_fr_poseidon_hash_chain": (
            E, [u64p, c_size, RawFr, RawFr, RawFr, RawFr, u64p, c_size, u64p, c_size]),
        "midnight_sppark_bls12_381_fr_poseidon_hash_chain_witness": (
            E, [u64p, c_size, RawFr, RawFr, RawFr, RawFr, u64p, c_size, u64p, c_size]),
        "midnight_sppark_bls12_381_fr_poseidon_advice_from_witness": (
            E, [vpp, u64p, c_size, c_size, u64p, c_size, ctypes.POINTER(ctypes.c_uint16), c_size]),
        "midnight_sppark_bls12_381_fr_lookup_permute_batch": (
            E, [vpp, u64p, u64p, c_size, c_size, c_size, vp, vp, c_size]),
        "midnight_sppark_bls12_381_trace_context_lookup_permute": (
            E, [vpp, vp, vp, c_size, vp, c_size, vp, c_size, RawFr, c_size, vp, vp, c_size]),
        "midnight_sppark_bls12_381_trace_context_trash": (
            E, [vpp, vp, vp, c_size, vp, c_size, vp, c_size, RawFr]),
        "midnight_sppark_bls12_381_trace_context_permutation": (
            E, [vpp, vp, vp, c_size, vp, c_size, RawFr, RawFr, u64p, c_size, vp, c_size, c_size]),
        "midnight_hip_plonk_keygen": (
            E, [vpp, u32, vp, c_size, u32, vp, c_size, c_size, vp, c_size, vp, c_size, vp, c_size, vp, RawGraph,
                ctypes.c_char_p]),
        "midnight_hip_plonk_keygen_with_arguments": (
            E, [vpp, u32, vp, c_size, u32, vp, c_size, c_size, vp, c_size, vp, c_size, vp, c_size, vp, RawGraph,
                ctypes.c_char_p, vp]),
        "midnight_hip_plonk_prove_with_instances": (
            E, [vp, u64p, c_size, vp, vp, c_size, ctypes.c_uint64, ctypes.c_char_p, c_size, ctypes.POINTER(c_size)]),
        "midnight_hip_plonk_key_free": (None, [vp]),
        "midnight_hip_plonk_vk": (E, [vp, u64p, u64p, u64p]),
        "midnight_hip_plonk_prove": (E, [vp, u64p, c_size, ctypes.c_uint64, ctypes.c_char_p, c_size,
                                         ctypes.POINTER(c_size)]),
        "midnight_hip_plonk_proof_size": (E, [vp, ctypes.POINTER(c_size)]),
        "midnight_hip_plonk_set_rng_threads": (E, [vp, u32]),
        "midnight_hip_plonk_set_graphs": (E, [vp, ctypes.c_bool]),
        "midnight_hip_srs_load": (E, [vpp, ctypes.c_char_p, c_size, u32]),
        "midnight_hip_srs_load_file": (E, [vpp, ctypes.c_char_p, u32]),
        "midnight_hip_srs_info": (E, [vp, ctypes.POINTER(u32), vpp, vpp]),
        "midnight_hip_srs_g2": (E, [vp, ctypes.c_char_p, c_size, ctypes.POINTER(c_size)]),
        "midnight_hip_srs_free": (None, [vp]),
        "midnight_hip_g1_msm_context_tuning": (E, [vp, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                                   ctypes.POINTER(ctypes.c_float)]),
        "midnight_sppark_test_mask_delay": (None, [u32]),
        "midnight_sppark_test_delayed_scale": (E, [vp, RawFr, u32]),
        "midnight_hip_fr_device_vec_ntt_zeta_noalloc": (E, [vp, u32, ctypes.c_bool, RawFr, RawFr, RawFr]),
        "midnight_hip_fr_sample_random_poly": (E, [vpp, u32, c_size, ctypes.c_char_p, c_size]),
        "midnight_hip_proof_server_create": (E, [vpp, u32, ctypes.c_uint64]),
        "midnight_hip_proof_server_submit": (E, [vp, vp, u64p, c_size, ctypes.c_uint64,
                                                 ctypes.POINTER(ctypes.c_uint64)]),
        "midnight_hip_proof_server_submit_with_instances": (
            E, [vp, vp, u64p, c_size, vp, vp, c_size, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
        "midnight_hip_proof_server_wait": (E, [vp, ctypes.c_uint64, ctypes.c_char_p, c_size, ctypes.POINTER(c_size),
                                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "midnight_hip_proof_server_free": (None, [vp]),
        "midnight_hip_test_msm_tree_levels": (None, [u32]),
        "midnight_hip_profile_enable": (None, [ctypes.c_int]),
        "midnight_hip_profile_read": (
            ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args


def check(err):
    if err.code != 0:
        msg = ""
        if err.message:
            msg = ctypes.string_at(err.message).decode(errors="replace")
            _libc.free(err.message)
        raise MidnightHipError(err.code, msg)


def P(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], "need C-contiguous uint64"
    return a.ctypes.data_as(u64p)


class RawDevicePolyRef(ctypes.Structure):
    """(source, index): source 0 dynamic, 1 fixed, 2 extra device vector."""
    _fields_ = [("source", ctypes.c_uint32), ("index", ctypes.c_uint32)]


def _u32(values):
    a = np.ascontiguousarray(values, dtype=np.uint32).reshape(-1)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _fr_array(values):
    a = np.ascontiguousarray(values, dtype=np.uint64)
    assert a.ndim == 2 and a.shape[1] == 4
    return a


# ---------------------------------------------------------------- lanes

def set_lane(lane):
    check(lib().midnight_sppark_execution_lane_set(lane))


# ---------------------------------------------------------------- NTT family

def fr_ntt(values, log_n, inverse=False):
    a = np.array(_fr_array(values), copy=True)
    assert a.shape[0] == 1 << log_n
    check(lib().midnight_sppark_bls12_381_fr_ntt(P(a), log_n, inverse))
    return a


def fr_ntt_zeta(values, log_n, inverse, scale, first, second):
    a = np.array(_fr_array(values), copy=True)
    check(lib().midnight_sppark_bls12_381_fr_ntt_zeta(P(a), log_n, inverse, RawFr.of(scale), RawFr.of(first),
                                                       RawFr.of(second)))
    return a


def fr_divide_ntt_zeta(values, log_n, factors, scale, first, second):
    a = np.array(_fr_array(values), copy=True)
    f = _fr_array(factors)
    check(lib().midnight_sppark_bls12_381_fr_divide_ntt_zeta(P(a), log_n, P(f), f.shape[0], RawFr.of(scale),
                                                              RawFr.of(first), RawFr.of(second)))
    return a


def fr_coeff_to_extended_batch(inputs, input_len, count, ext_log_n, first, second):
    a = _fr_array(inputs)
    assert a.shape[0] == input_len * count
    out = np.zeros(((1 << ext_log_n) * count, 4), dtype=np.uint64)
    check(lib().midnight_sppark_bls12_381_fr_coeff_to_extended_batch(
        P(out), P(a), input_len, count, ext_log_n, RawFr.of(first), RawFr.of(second)))
    return out


def fr_mul_assign(values, scalar):
    a = np.array(_fr_array(values), copy=True)
    check(lib().midnight_sppark_bls12_381_fr_mul_assign(P(a), RawFr.of(scalar), a.shape[0]))
    return a


def fr_mul_assign_zeta(values, first, second):
    a = np.array(_fr_array(values), copy=True)
    check(lib().midnight_sppark_bls12_381_fr_mul_assign_zeta(P(a), RawFr.of(first), RawFr.of(second), a.shape[0]))
    return a


def fr_mul_assign_repeating(values, factors):
    a = np.array(_fr_array(values), copy=True)
    f = _fr_array(factors)
    check(lib().midnight_sppark_bls12_381_fr_mul_assign_repeating(P(a), P(f), f.shape[0], a.shape[0]))
    return a


# ---------------------------------------------------------------- device vectors

class DeviceVec:
    """Owning handle of a resident Fr vector (scalar_device_vec_t)."""

    def __init__(self, handle):
        self.handle = ctypes.c_void_p(handle)

    @classmethod
    def upload(cls, values):
        a = _fr_array(values)
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_fr_device_vec_upload(ctypes.byref(h), P(a), a.shape[0]))
        return cls(h.value)

    @classmethod
    def upload_many(cls, arrays):
        arrs = [_fr_array(x) for x in arrays]
        ptrs = (u64p * len(arrs))(*[P(x) for x in arrs])
        lens = (c_size * len(arrs))(*[x.shape[0] for x in arrs])
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_fr_device_vec_upload_many(ctypes.byref(h), ptrs, lens, len(arrs)))
        return cls(h.value)

    def __len__(self):
        return lib().midnight_hip_fr_device_vec_len(self.handle)

    def download(self):
        n = len(self)
        out = np.zeros((n, 4), dtype=np.uint64)
        check(lib().midnight_sppark_bls12_381_fr_device_vec_download(P(out), self.handle, n))
        return out

    def free(self):
        if self.handle and self.handle.value:
            lib().midnight_sppark_bls12_381_fr_device_vec_free(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def ntt_zeta(self, log_n, inverse, scale, first, second):
        check(lib().midnight_sppark_bls12_381_fr_device_vec_ntt_zeta(
            self.handle, log_n, inverse, RawFr.of(scale), RawFr.of(first), RawFr.of(second)))

    def ntt_zeta_noalloc(self, log_n, inverse, scale, first, second):
        check(lib().midnight_hip_fr_device_vec_ntt_zeta_noalloc(
            self.handle, log_n, inverse, RawFr.of(scale), RawFr.of(first), RawFr.of(second)))

    def ntt_zeta_batch(self, log_n, count, inverse, scale, first, second):
        check(lib().midnight_sppark_bls12_381_fr_device_vec_ntt_zeta_batch(
            self.handle, log_n, count, inverse, RawFr.of(scale), RawFr.of(first), RawFr.of(second)))

    def divide_ntt_zeta(self, log_n, factors, scale, first, second):
        f = _fr_array(factors)
        check(lib().midnight_sppark_bls12_381_fr_device_vec_divide_ntt_zeta(
            self.handle, log_n, P(f), f.shape[0], RawFr.of(scale), RawFr.of(first), RawFr.of(second)))

    def ntt_multi(self, offsets, log_ns, inverse):
        n = len(offsets)
        offs = (c_size * n)(*offsets)
        logs = (ctypes.c_uint32 * n)(*log_ns)
        check(lib().midnight_hip_fr_device_vec_ntt_multi(self.handle, offs, logs, n, inverse))

    def scatter_host(self, indices, values):
        idx, pidx = _u32(indices)
        v = _fr_array(values) if len(idx) else np.zeros((0, 4), dtype=np.uint64)
        check(lib().midnight_sppark_bls12_381_fr_device_vec_scatter_host(self.handle, pidx, _ptr(v), len(idx)))

    def concat_host(self, suffix):
        sfx = _fr_array(suffix) if len(suffix) else None
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_fr_device_vec_concat_host(
            ctypes.byref(h), self.handle, _ptr(sfx), 0 if sfx is None else sfx.shape[0]))
        return DeviceVec(h.value)

    @classmethod
    def concat_many_host(cls, vecs, suffix):
        arr = (ctypes.c_void_p * len(vecs))(*[v.handle.value for v in vecs])
        sfx = _fr_array(suffix) if len(suffix) else None
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_fr_device_vec_concat_many_host(
            ctypes.byref(h), ctypes.cast(arr, ctypes.POINTER(ctypes.c_void_p)), len(vecs), _ptr(sfx),
            0 if sfx is None else sfx.shape[0]))
        return cls(h.value)

    def gather_polys(self, poly_len, indices):
        idx, pidx = _u32(indices)
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_fr_device_vec_gather_polys(
            ctypes.byref(h), self.handle, poly_len, pidx, len(idx)))
        return DeviceVec(h.value)

    def download_nonzero_suffixes(self, scalar_offset, poly_len, count):
        """(concatenated suffixes, starts, nonzero counts)"""
        starts = (c_size * count)()
        counts = (c_size * count)()
        total = c_size()
        check(lib().midnight_sppark_bls12_381_fr_device_vec_download_nonzero_suffixes(
            None, 0, ctypes.byref(total), starts, counts, self.handle, scalar_offset, poly_len, count))
        out = np.zeros((max(total.value, 1), 4), dtype=np.uint64)
        check(lib().midnight_sppark_bls12_381_fr_device_vec_download_nonzero_suffixes(
            P(out), out.shape[0], ctypes.byref(total), starts, counts, self.handle, scalar_offset, poly_len, count))
        return out[:total.value], list(starts), list(counts)

    def poly_lincomb(self, poly_len, poly_count, scalars):
        sc = _fr_array(scalars)
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_fr_device_poly_lincomb(
            ctypes.byref(h), self.handle, poly_len, poly_count, P(sc), sc.shape[0]))
        return DeviceVec(h.value)

    def eval_polys(self, poly_len, poly_count, point):
        out = np.zeros((poly_count, 4), dtype=np.uint64)
        check(lib().midnight_sppark_bls12_381_fr_device_eval_polys(P(out), self.handle, poly_len, poly_count,
                                                                   RawFr.of(point)))
        return out

    def quotient_pieces(self, poly_len, piece_count, blinds):
        bl = _fr_array(blinds)
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_fr_device_quotient_pieces(
            ctypes.byref(h), self.handle, poly_len, piece_count, P(bl), bl.shape[0]))
        return DeviceVec(h.value)


def device_coeff_to_extended_batch(inputs, input_len, count, ext_log_n, first, second):
    a = _fr_array(inputs)
    h = ctypes.c_void_p()
    check(lib().midnight_sppark_bls12_381_fr_device_coeff_to_extended_batch(
        ctypes.byref(h), P(a), input_len, count, ext_log_n, RawFr.of(first), RawFr.of(second)))
    return DeviceVec(h.value)


def device_coeff_to_extended_from_device_batch(src, input_len, count, ext_log_n, first, second):
    h = ctypes.c_void_p()
    check(lib().midnight_sppark_bls12_381_fr_device_coeff_to_extended_from_device_batch(
        ctypes.byref(h), src.handle, input_len, count, ext_log_n, RawFr.of(first), RawFr.of(second)))
    return DeviceVec(h.value)


def test_mask_delay(ms):
    """delay (ms) of the calling thread's scrambled-batch mask producer"""
    lib().midnight_sppark_test_mask_delay(ms)


def test_delayed_scale(dvec, scale, ms):
    """scale a device vector behind a delay on the lane's stream; returns at once"""
    check(lib().midnight_sppark_test_delayed_scale(dvec.handle, RawFr.of(scale), ms))


# ---------------------------------------------------------------- MSM

def _aff_array(points):
    a = np.ascontiguousarray(points, dtype=np.uint64)
    assert a.ndim == 2 and a.shape[1] == 12
    return a


def g1_msm(points, scalars):
    p = _aff_array(points)
    s = _fr_array(scalars)
    out = np.zeros(18, dtype=np.uint64)
    check(lib().midnight_sppark_bls12_381_g1_msm(P(out), P(p), P(s), p.shape[0]))
    return out


def g1_msm_batch(points, scalars, nmsms):
    p = _aff_array(points)
    s = _fr_array(scalars)
    out = np.zeros((nmsms, 18), dtype=np.uint64)
    check(lib().midnight_sppark_bls12_381_g1_msm_batch(P(out), P(p), P(s), p.shape[0], nmsms))
    return out


def default_config(npoints):
    """Mirror of default_msm_config (bls12_381_g1_msm.cu:70-83)."""
    k = 0
    cap = npoints - 1 if npoints > 1 else 0
    while cap:
        k += 1
        cap >>= 1
    wb = 0 if k == 0 else 10
    if 16 <= k <= 18:
        wb = 12
    if k > 18:
        wb = 0
    return RawMsmConfig(k, 100000, 32768, wb, 3, 2, 1, 4, 0, 1 if 14 <= k <= 18 else 0)


class MsmContext:
    """Resident point set (msm_context_t)."""

    def __init__(self, points, config=None, _handle=None, _npoints=0, _owned=True):
        self.owned = _owned
        if _handle is not None:
            self.handle = _handle
            self.npoints = _npoints
            return
        p = _aff_array(points)
        self.npoints = p.shape[0]
        h = ctypes.c_void_p()
        if config is None:
            check(lib().midnight_sppark_bls12_381_g1_msm_context_create(ctypes.byref(h), P(p), p.shape[0]))
        else:
            check(lib().midnight_sppark_bls12_381_g1_msm_context_create_with_config(
                ctypes.byref(h), P(p), p.shape[0], config))
        self.handle = h

    def free(self):
        if self.handle and self.handle.value and self.owned:
            lib().midnight_sppark_bls12_381_g1_msm_context_free(self.handle)
        self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def batch(self, scalars, npoints, nmsms):
        s = _fr_array(scalars)
        out = np.zeros((nmsms, 18), dtype=np.uint64)
        check(lib().midnight_sppark_bls12_381_g1_msm_context_batch(P(out), self.handle, P(s), npoints, nmsms))
        return out

    def device_batch(self, dvec, scalar_offset, npoints, nmsms):
        out = np.zeros((nmsms, 18), dtype=np.uint64)
        check(lib().midnight_sppark_bls12_381_g1_msm_context_device_batch(
            P(out), self.handle, dvec.handle, scalar_offset, npoints, nmsms))
        return out

    def device_ragged(self, dvec, slices):
        n = len(slices)
        arr = (RawMsmSlice * n)(*[RawMsmSlice(*s) for s in slices])
        out = np.zeros((n, 18), dtype=np.uint64)
        check(lib().midnight_sppark_bls12_381_g1_msm_context_device_ragged(P(out), self.handle, dvec.handle, arr, n))
        return out

    def device_scrambled_batch(self, dvec, scalar_offset, npoints, nmsms, mask):
        out = np.zeros((nmsms, 18), dtype=np.uint64)
        check(lib().midnight_sppark_bls12_381_g1_msm_context_device_scrambled_batch(
            P(out), self.handle, dvec.handle, scalar_offset, npoints, nmsms, mask.handle))
        return out

    @classmethod
    def from_multiples(cls, dvec):
        """Context over the points [t_i]G, t_i the (Montgomery) values of dvec,
        computed on the device."""
        h = ctypes.c_void_p()
        check(lib().midnight_hip_g1_msm_context_create_from_multiples(ctypes.byref(h), dvec.handle))
        return cls(None, _handle=h, _npoints=len(dvec))

    def points(self, n=None):
        n = self.npoints if n is None else n
        out = np.zeros((n, 12), dtype=np.uint64)
        check(lib().midnight_hip_g1_msm_context_points_download(P(out), self.handle, n))
        return out

    def tuning(self):
        """(table or plain window, chunk length, autotuned ms)"""
        w, c, ms = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_float()
        check(lib().midnight_hip_g1_msm_context_tuning(self.handle, ctypes.byref(w), ctypes.byref(c),
                                                       ctypes.byref(ms)))
        return w.value, c.value, ms.value

    def last_stats(self):
        e = ctypes.c_uint64()
        w = ctypes.c_uint32()
        check(lib().midnight_hip_g1_msm_context_last_stats(self.handle, ctypes.byref(e), ctypes.byref(w)))
        return e.value, w.value

    def device_multi(self, dvec, scalar_offsets, sizes):
        n = len(sizes)
        offs = (c_size * n)(*scalar_offsets)
        sz = (c_size * n)(*sizes)
        out = np.zeros((n, 18), dtype=np.uint64)
        check(lib().midnight_hip_g1_msm_context_device_multi(P(out), self.handle, dvec.handle, offs, sz, n))
        return out


class BatchPlan:
    """Batched iNTT + MSM plan over a resident context and device vector:
    instance i = inverse NTT (in place) of the 2^log_ns[i] values at
    offsets[i], then their MSM against points [0, 2^log_ns[i])."""

    GRAPH = 1
    AUTOTUNE = 2

    def __init__(self, ctx, dvec, offsets, log_ns, flags=GRAPH | AUTOTUNE):
        n = len(offsets)
        self.count = n
        offs = (c_size * n)(*offsets)
        logs = (ctypes.c_uint32 * n)(*log_ns)
        h = ctypes.c_void_p()
        check(lib().midnight_hip_batch_plan_create(ctypes.byref(h), ctx.handle, dvec.handle, offs, logs, n, flags))
        self.handle = h
        self._refs = (ctx, dvec)   # the plan refers to both

    def run(self):
        out = np.zeros((self.count, 18), dtype=np.uint64)
        check(lib().midnight_hip_batch_plan_run(self.handle, P(out)))
        return out

    def info(self):
        w, L, pc, g = (ctypes.c_uint32() for _ in range(4))
        ms = ctypes.c_float()
        check(lib().midnight_hip_batch_plan_info(self.handle, ctypes.byref(w), ctypes.byref(L), ctypes.byref(pc),
                                                 ctypes.byref(g), ctypes.byref(ms)))
        return {"window_bits": w.value, "chunk_len": L.value, "precomputed": bool(pc.value), "graph": bool(g.value),
                "tuned_msm_ms": round(ms.value, 3)}

    def free(self):
        if self.handle and self.handle.value:
            lib().midnight_hip_batch_plan_free(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- diagnostics

def to_montgomery(dvec):
    """In place: canonical values -> Montgomery form (a -> a R mod r)."""
    check(lib().midnight_hip_fr_device_vec_to_montgomery(dvec.handle))


def profile_enable(on=True):
    lib().midnight_hip_profile_enable(1 if on else 0)


def profile_read(name):
    ms = ctypes.c_double()
    cnt = ctypes.c_uint64()
    lib().midnight_hip_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
    return ms.value, cnt.value


def valu_peak():
    r = ctypes.c_double()
    clk = ctypes.c_double()
    cus = ctypes.c_int()
    check(lib().midnight_hip_valu_peak(ctypes.byref(r), ctypes.byref(clk), ctypes.byref(cus)))
    return r.value, clk.value, cus.value


def copy_bandwidth(nbytes):
    """read + write bytes/s of one device copy of nbytes by a plain kernel"""
    r = ctypes.c_double()
    check(lib().midnight_hip_copy_bandwidth(nbytes, ctypes.byref(r)))
    return r.value


def fr29_product_latency():
    """seconds per dependent radix-2^29 Fr product on one wave"""
    t = ctypes.c_double()
    check(lib().midnight_hip_fr29_product_latency(ctypes.byref(t)))
    return t.value


def field_mul_rate(blocks=0):
    fp = ctypes.c_double()
    fr = ctypes.c_double()
    check(lib().midnight_hip_field_mul_rate(blocks, ctypes.byref(fp), ctypes.byref(fr)))
    return fp.value, fr.value


# ---------------------------------------------------------------- quotient numerator

def _ptr(a):
    """Data pointer of an optional numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _cols(a, n):
    """(cols * n, 4) or (cols, n, 4) Fr array -> (contiguous array, cols)."""
    if a is None:
        return None, 0
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4)
    if n == 0:
        return a, 0
    assert a.shape[0] % n == 0, "column array length is not a multiple of the row count"
    return a, a.shape[0] // n


def _graphs(gs):
    gs = list(gs)
    if not gs:
        return None, None
    arr = (RawGraph * len(gs))(*[g.as_raw() for g in gs])
    return arr, ctypes.addressof(arr)


def _column_refs(refs):
    refs = list(refs)
    if not refs:
        return None, None
    arr = (RawColumnRef * len(refs))(*[RawColumnRef(k, i) for k, i in refs])
    return arr, ctypes.addressof(arr)


def eval_numerator(n, *, beta, gamma, theta, trash_challenge, y, omega, delta_start, delta_const,
                   rot_scale, last_rotation, chunk_len, custom_graph=None, run_custom=True,
                   initial_out=None, fixed=None, advice=None, instance=None, challenges=None,
                   l0=None, l_last=None, l_active=None, permutation_products=None,
                   permutation_columns=(), permutation_pk_cosets=None, lookup_product_cosets=None,
                   lookup_input_cosets=None, lookup_table_cosets=None, lookup_graphs=(),
                   trash_cosets=None, trash_selector_fixed_indices=(), trash_graphs=()):
    """midnight_sppark_bls12_381_eval_numerator on host arrays.  Column
    arrays are (cols * n, 4) uint64 Montgomery Fr; graphs are
    numerator_graph.RawGraphBuffers; permutation_columns are (kind, index)
    pairs.  Returns the (n, 4) numerator values."""
    out = np.zeros((n, 4), dtype=np.uint64)
    fixed, nf = _cols(fixed, n)
    advice, na = _cols(advice, n)
    instance, ni = _cols(instance, n)
    ch = None if challenges is None else np.ascontiguousarray(challenges, dtype=np.uint64).reshape(-1, 4)
    init = None if initial_out is None else np.ascontiguousarray(initial_out, dtype=np.uint64).reshape(-1, 4)
    lag = [None if x is None else np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 4)
           for x in (l0, l_last, l_active)]
    pp, nsets = _cols(permutation_products, n)
    pc, ncos = _cols(permutation_pk_cosets, n)
    lp, nlook = _cols(lookup_product_cosets, n)
    li, _ = _cols(lookup_input_cosets, n)
    lt, _ = _cols(lookup_table_cosets, n)
    tc, ntrash = _cols(trash_cosets, n)
    col_arr, col_ptr = _column_refs(permutation_columns)
    sel = np.ascontiguousarray(trash_selector_fixed_indices, dtype=np.uint32) if ntrash else None
    lg_arr, lg_ptr = _graphs(lookup_graphs)
    tg_arr, tg_ptr = _graphs(trash_graphs)
    cg = custom_graph.as_raw() if custom_graph is not None else empty_raw_graph()
    check(lib().midnight_sppark_bls12_381_eval_numerator(
        P(out), n, _ptr(init), 1 if run_custom else 0, _ptr(fixed), nf, _ptr(advice), na, _ptr(instance), ni,
        _ptr(ch), 0 if ch is None else ch.shape[0],
        RawFr.of(beta), RawFr.of(gamma), RawFr.of(theta), RawFr.of(trash_challenge), RawFr.of(y),
        RawFr.of(omega), RawFr.of(delta_start), RawFr.of(delta_const), rot_scale, last_rotation, chunk_len,
        _ptr(lag[0]), _ptr(lag[1]), _ptr(lag[2]), _ptr(pp), nsets, col_ptr, len(permutation_columns),
        _ptr(pc), ncos, _ptr(lp), _ptr(li), _ptr(lt), nlook, _ptr(tc), _ptr(sel), ntrash, cg, lg_ptr, tg_ptr))
    return out


def eval_graph_batch(n, graphs, *, beta, gamma, theta, trash_challenge, y, rot_scale, fixed=None,
                     advice=None, instance=None, challenges=None):
    """midnight_sppark_bls12_381_eval_graph_batch: (len(graphs), n, 4)."""
    fixed, nf = _cols(fixed, n)
    advice, na = _cols(advice, n)
    instance, ni = _cols(instance, n)
    ch = None if challenges is None else np.ascontiguousarray(challenges, dtype=np.uint64).reshape(-1, 4)
    g_arr, g_ptr = _graphs(graphs)
    out = np.zeros((len(graphs) * n, 4), dtype=np.uint64)
    check(lib().midnight_sppark_bls12_381_eval_graph_batch(
        P(out), n, _ptr(fixed), nf, _ptr(advice), na, _ptr(instance), ni, _ptr(ch),
        0 if ch is None else ch.shape[0], RawFr.of(beta), RawFr.of(gamma), RawFr.of(theta),
        RawFr.of(trash_challenge), RawFr.of(y), rot_scale, g_ptr, len(graphs)))
    return out.reshape(len(graphs), n, 4)


class NumeratorContext:
    """Resident key data for repeated numerator evaluations
    (numerator_context_create / _eval / _free)."""

    def __init__(self, n, input_len, *, fixed, fixed_lagrange, l0, l_last, l_active, permutation_columns,
                 permutation_pk_cosets, permutation_pk_lagrange, custom_graph, omega, trace_omega, g_coset,
                 delta_const, rot_scale, last_rotation, chunk_len, lookup_graphs=(), lookup_input_graphs=(),
                 lookup_table_graphs=(), trash_graphs=(), trash_selector_fixed_indices=()):
        fixed, nf = _cols(fixed, n)
        fl, _ = _cols(fixed_lagrange, input_len)
        pc, ncos = _cols(permutation_pk_cosets, n)
        pl, _ = _cols(permutation_pk_lagrange, input_len)
        lag = [np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 4) for x in (l0, l_last, l_active)]
        col_arr, col_ptr = _column_refs(permutation_columns)
        ntrash = len(trash_graphs)
        sel = np.ascontiguousarray(trash_selector_fixed_indices, dtype=np.uint32) if ntrash else None
        lg = _graphs(lookup_graphs)
        lig = _graphs(lookup_input_graphs)
        ltg = _graphs(lookup_table_graphs)
        tg = _graphs(trash_graphs)
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_numerator_context_create(
            ctypes.byref(h), n, input_len, _ptr(fixed), _ptr(fl), nf, _ptr(lag[0]), _ptr(lag[1]), _ptr(lag[2]),
            col_ptr, len(permutation_columns), _ptr(pc), _ptr(pl), ncos, _ptr(sel), ntrash,
            custom_graph.as_raw(), lg[1], lig[1], ltg[1], len(lookup_graphs), tg[1],
            RawFr.of(omega), RawFr.of(trace_omega), RawFr.of(g_coset), RawFr.of(delta_const),
            rot_scale, last_rotation, chunk_len))
        self.handle = h
        self.n = n
        self.lookup_count = len(lookup_graphs)

    def eval(self, columns, total_columns, *, beta, gamma, theta, trash_challenge, y, advice_offset=0,
             advice_cols=0, instance_offset=0, instance_cols=0, challenges=None, permutation_offset=0,
             permutation_sets=0, lookup_product_offset=0, lookup_input_offset=0, lookup_table_offset=0,
             lookup_count=0, trash_offset=0, trash_count=0):
        """Numerator from a DeviceVec of total_columns * n resident values;
        returns a new DeviceVec of n values."""
        ch = None if challenges is None else np.ascontiguousarray(challenges, dtype=np.uint64).reshape(-1, 4)
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_numerator_context_eval(
            ctypes.byref(h), self.handle, columns.handle if columns is not None else None, total_columns,
            advice_offset, advice_cols, instance_offset, instance_cols, _ptr(ch),
            0 if ch is None else ch.shape[0], RawFr.of(beta), RawFr.of(gamma), RawFr.of(theta),
            RawFr.of(trash_challenge), RawFr.of(y), permutation_offset, permutation_sets,
            lookup_product_offset, lookup_input_offset, lookup_table_offset, lookup_count,
            trash_offset, trash_count))
        return DeviceVec(h.value)

    def trace_permutation(self, advice, advice_cols, instance, instance_cols, *, beta, gamma, delta_powers,
                          blinds, blinding_factors):
        """trace_context_permutation: permutation products from resident
        Lagrange advice / instance (DeviceVec or None); returns a DeviceVec."""
        dp = _fr_array(delta_powers)
        bl = None if blinds is None or len(blinds) == 0 else _fr_array(blinds)
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_trace_context_permutation(
            ctypes.byref(h), self.handle, advice.handle if advice is not None else None, advice_cols,
            instance.handle if instance is not None else None, instance_cols, RawFr.of(beta), RawFr.of(gamma),
            P(dp), dp.shape[0], _ptr(bl), 0 if bl is None else bl.shape[0], blinding_factors))
        return DeviceVec(h.value)

    def trace_lookup_permute(self, advice, advice_cols, instance, instance_cols, challenges, theta, usable_rows,
                             input_blinds, table_blinds):
        """trace_context_lookup_permute: DeviceVec [compressed inputs | tables | permuted inputs | tables]."""
        ch = None if challenges is None or len(challenges) == 0 else _fr_array(challenges)
        ib = _blinds_or_empty(input_blinds)
        tb = _blinds_or_empty(table_blinds)
        blind_count = np.ascontiguousarray(input_blinds, dtype=np.uint64).reshape(-1, 4).shape[0]
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_trace_context_lookup_permute(
            ctypes.byref(h), self.handle, advice.handle if advice is not None else None, advice_cols,
            instance.handle if instance is not None else None, instance_cols, _ptr(ch),
            0 if ch is None else ch.shape[0], RawFr.of(theta), usable_rows, _ptr(ib), _ptr(tb),
            blind_count // max(self.lookup_count, 1)))
        return DeviceVec(h.value)

    def trace_trash(self, advice, advice_cols, instance, instance_cols, challenges, trash_challenge):
        """trace_context_trash: the trash graphs on the Lagrange domain; a DeviceVec."""
        ch = None if challenges is None or len(challenges) == 0 else _fr_array(challenges)
        h = ctypes.c_void_p()
        check(lib().midnight_sppark_bls12_381_trace_context_trash(
            ctypes.byref(h), self.handle, advice.handle if advice is not None else None, advice_cols,
            instance.handle if instance is not None else None, instance_cols, _ptr(ch),
            0 if ch is None else ch.shape[0], RawFr.of(trash_challenge)))
        return DeviceVec(h.value)

    def free(self):
        if self.handle and self.handle.value:
            lib().midnight_sppark_bls12_381_numerator_context_free(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- grand products

def fr_shifted_prefix_product(values, initial):
    """out[0] = initial, out[i] = initial * prod_{j<i} values[j]"""
    a = np.array(_fr_array(values), copy=True)
    check(lib().midnight_sppark_bls12_381_fr_shifted_prefix_product(P(a), RawFr.of(initial), a.shape[0]))
    return a


def fr_permutation_products(columns, permutations, n, chunk_len, *, beta, gamma, omega, delta_powers, blinds,
                            blinding_factors):
    """Permutation product columns (sets * n values) as a DeviceVec."""
    cols = _fr_array(columns)
    perm = _fr_array(permutations)
    dp = _fr_array(delta_powers)
    bl = None if blinding_factors == 0 else _fr_array(blinds)
    h = ctypes.c_void_p()
    check(lib().midnight_sppark_bls12_381_fr_permutation_products(
        ctypes.byref(h), P(cols), P(perm), n, cols.shape[0] // n, chunk_len, RawFr.of(beta), RawFr.of(gamma),
        RawFr.of(omega), P(dp), _ptr(bl), blinding_factors))
    return DeviceVec(h.value)


def fr_sort_canonical(values):
    a = np.array(_fr_array(values), copy=True)
    check(lib().midnight_sppark_bls12_381_fr_sort_canonical(P(a), a.shape[0]))
    return a


POSEIDON_HASH_WITNESS_LEN = 256


def _poseidon_chain(fn, per_hash, length, initial, tail0, tail1, input_len, mds, round_constants):
    m = _fr_array(mds)
    rc = _fr_array(round_constants)
    out = np.zeros((max(length * per_hash, 1), 4), dtype=np.uint64)
    check(fn(P(out), length, RawFr.of(initial), RawFr.of(tail0), RawFr.of(tail1), RawFr.of(input_len),
             P(m), m.shape[0], P(rc), rc.shape[0]))
    return out[:length * per_hash]


def fr_poseidon_hash_chain(length, initial, tail0, tail1, input_len, mds, round_constants):
    """(length, 4) chain outputs: hash i permutes [previous, tail0, input_len], adds tail1 to lane 0
    and permutes again; previous starts at `initial`."""
    return _poseidon_chain(lib().midnight_sppark_bls12_381_fr_poseidon_hash_chain, 1, length, initial, tail0,
                           tail1, input_len, mds, round_constants)


def fr_poseidon_hash_chain_witness(length, initial, tail0, tail1, input_len, mds, round_constants):
    """(length * 256, 4) circuit witness of the same chain"""
    return _poseidon_chain(lib().midnight_sppark_bls12_381_fr_poseidon_hash_chain_witness,
                           POSEIDON_HASH_WITNESS_LEN, length, initial, tail0, tail1, input_len, mds,
                           round_constants)


def fr_poseidon_advice_from_witness(witness, hash_count, domain_len, metadata, layout):
    """DeviceVec of 8 * domain_len advice values (column-major) built from a chain witness"""
    w = _fr_array(witness)
    meta = _fr_array(metadata)
    lay = np.ascontiguousarray(layout, dtype=np.uint16).reshape(-1)
    h = ctypes.c_void_p()
    check(lib().midnight_sppark_bls12_381_fr_poseidon_advice_from_witness(
        ctypes.byref(h), P(w), hash_count, domain_len, P(meta), meta.shape[0],
        lay.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), lay.shape[0]))
    return DeviceVec(h.value)


def _blinds_or_empty(b):
    a = np.ascontiguousarray(b, dtype=np.uint64).reshape(-1, 4)
    return a if a.shape[0] else np.zeros((1, 4), dtype=np.uint64)


def fr_lookup_permute_batch(inputs, tables, n, count, usable_rows, input_blinds, table_blinds):
    """DeviceVec [compressed inputs | tables | permuted inputs | permuted tables]"""
    ci = _fr_array(inputs)
    ct = _fr_array(tables)
    ib = _blinds_or_empty(input_blinds)
    tb = _blinds_or_empty(table_blinds)
    h = ctypes.c_void_p()
    check(lib().midnight_sppark_bls12_381_fr_lookup_permute_batch(
        ctypes.byref(h), P(ci), P(ct), n, count, usable_rows, _ptr(ib), _ptr(tb), n - usable_rows))
    return DeviceVec(h.value)


def fr_invert_assign(values):
    a = np.array(_fr_array(values), copy=True)
    check(lib().midnight_sppark_bls12_381_fr_invert_assign(P(a), a.shape[0]))
    return a


def fr_lookup_products(ci, ct, pi, pt, n, count, *, beta, gamma, blinds, blinding_factors):
    """Lookup product columns (count * n values) as a DeviceVec."""
    arrs = [_fr_array(x) for x in (ci, ct, pi, pt)]
    bl = None if blinding_factors == 0 else _fr_array(blinds)
    h = ctypes.c_void_p()
    check(lib().midnight_sppark_bls12_381_fr_lookup_products(
        ctypes.byref(h), *[P(a) for a in arrs], n, count, RawFr.of(beta), RawFr.of(gamma), _ptr(bl),
        blinding_factors))
    return DeviceVec(h.value)


def fr_lookup_products_device(trace, n, count, *, beta, gamma, blinds, blinding_factors):
    bl = None if blinding_factors == 0 else _fr_array(blinds)
    h = ctypes.c_void_p()
    check(lib().midnight_sppark_bls12_381_fr_lookup_products_device(
        ctypes.byref(h), trace.handle, n, count, RawFr.of(beta), RawFr.of(gamma), _ptr(bl), blinding_factors))
    return DeviceVec(h.value)


def fr_poly_lincomb(polys, poly_len, scalars):
    pl = _fr_array(polys)
    sc = _fr_array(scalars)
    out = np.zeros((poly_len, 4), dtype=np.uint64)
    check(lib().midnight_sppark_bls12_381_fr_poly_lincomb(P(out), P(pl), poly_len, sc.shape[0], P(sc)))
    return out


def fr_multiopen_prepare(polys, poly_len, q_offsets, q_poly_indices, q_coefficients, points, query_poly_indices,
                         query_point_indices):
    """(q polynomials (q_count * poly_len), evaluations) from host polynomials."""
    pl = _fr_array(polys)
    qo, pqo = _u32(q_offsets)
    qi, pqi = _u32(q_poly_indices)
    qc = _fr_array(q_coefficients) if len(q_coefficients) else np.zeros((1, 4), dtype=np.uint64)
    pts = _fr_array(points)
    qp, pqp = _u32(query_poly_indices)
    qx, pqx = _u32(query_point_indices)
    q_count = len(qo) - 1
    q_out = np.zeros((q_count * poly_len, 4), dtype=np.uint64)
    ev = np.zeros((max(len(qp), 1), 4), dtype=np.uint64)
    check(lib().midnight_sppark_bls12_381_fr_multiopen_prepare(
        P(q_out), P(ev), P(pl), poly_len, pl.shape[0] // poly_len, pqo, len(qo), pqi, P(qc), len(qi), P(pts),
        pts.shape[0], pqp, pqx, len(qp)))
    return q_out, ev[:len(qp)]


def fr_multiopen_prepare_device(dynamic, fixed, extra, poly_len, poly_refs, q_offsets, q_poly_indices,
                                q_coefficients, points, query_poly_indices, query_point_indices, host_q=True):
    """(q polynomials on the host or None, q DeviceVec, evaluations) from resident polynomials;
    poly_refs: (source, index) pairs."""
    refs = (RawDevicePolyRef * len(poly_refs))(*[RawDevicePolyRef(a, b) for a, b in poly_refs])
    qo, pqo = _u32(q_offsets)
    qi, pqi = _u32(q_poly_indices)
    qc = _fr_array(q_coefficients) if len(q_coefficients) else np.zeros((1, 4), dtype=np.uint64)
    pts = _fr_array(points)
    qp, pqp = _u32(query_poly_indices)
    qx, pqx = _u32(query_point_indices)
    q_count = len(qo) - 1
    q_out = np.zeros((q_count * poly_len, 4), dtype=np.uint64) if host_q else None
    ev = np.zeros((max(len(qp), 1), 4), dtype=np.uint64)
    h = ctypes.c_void_p()
    check(lib().midnight_sppark_bls12_381_fr_multiopen_prepare_device(
        _ptr(q_out), ctypes.byref(h), P(ev), dynamic.handle, fixed.handle, extra.handle, poly_len, refs,
        len(poly_refs), pqo, len(qo), pqi, P(qc), len(qi), P(pts), pts.shape[0], pqp, pqx, len(qp)))
    return q_out, DeviceVec(h.value), ev[:len(qp)]


def fr_multiopen_f_poly_device(q_vec, poly_len, point_offsets, points, scalars):
    po, ppo = _u32(point_offsets)
    pts = _fr_array(points)
    sc = _fr_array(scalars)
    h = ctypes.c_void_p()
    check(lib().midnight_sppark_bls12_381_fr_multiopen_f_poly_device(
        ctypes.byref(h), q_vec.handle, poly_len, ppo, len(po), P(pts), pts.shape[0], P(sc), sc.shape[0]))
    return DeviceVec(h.value)


def fr_multiopen_pi_device(q_vec, f_vec, poly_len, scalars, point):
    sc = _fr_array(scalars)
    h = ctypes.c_void_p()
    check(lib().midnight_sppark_bls12_381_fr_multiopen_pi_device(
        ctypes.byref(h), q_vec.handle, f_vec.handle, poly_len, P(sc), sc.shape[0], RawFr.of(point)))
    return DeviceVec(h.value)


def fr_sample_random_poly(log_n, threads, seeds):
    """device vector of 2^log_n random Fr, one ChaCha20Rng seed (32 bytes of
    `seeds`) per parallelize chunk of `threads` workers"""
    h = ctypes.c_void_p()
    check(lib().midnight_hip_fr_sample_random_poly(ctypes.byref(h), log_n, threads, bytes(seeds),
                                                   len(seeds) // 32))
    return DeviceVec(h.value)


class Srs:
    """KZG parameters (bls_midnight_2pK files) resident as two MSM contexts:
    g ([tau^i]G) and g_lagrange.  format: 0 Processed, 1 RawBytes,
    2 RawBytesUnchecked."""

    PROCESSED, RAW_BYTES, RAW_BYTES_UNCHECKED = 0, 1, 2

    def __init__(self, data=None, path=None, format=1):
        h = ctypes.c_void_p()
        if path is not None:
            check(lib().midnight_hip_srs_load_file(ctypes.byref(h), os.fsencode(path), format))
        else:
            data = bytes(data)
            check(lib().midnight_hip_srs_load(ctypes.byref(h), data, len(data), format))
        self.handle = h
        k = ctypes.c_uint32()
        g, gl = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().midnight_hip_srs_info(h, ctypes.byref(k), ctypes.byref(g), ctypes.byref(gl)))
        self.k = k.value
        self.g = MsmContext(None, _handle=g, _npoints=1 << self.k, _owned=False)
        self.g_lagrange = MsmContext(None, _handle=gl, _npoints=1 << self.k, _owned=False)

    def g2(self):
        ln = c_size()
        buf = ctypes.create_string_buffer(1024)
        check(lib().midnight_hip_srs_g2(self.handle, buf, 1024, ctypes.byref(ln)))
        return buf.raw[:ln.value]

    def free(self):
        if self.handle and self.handle.value:
            lib().midnight_hip_srs_free(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- proof pipeline

class RawPlonkArguments(ctypes.Structure):
    """include/midnight_hip.h RawPlonkArguments"""
    _fields_ = [("instance_cols", ctypes.c_size_t), ("committed_instances", ctypes.c_size_t),
                ("instance_queries", ctypes.c_void_p), ("instance_query_count", ctypes.c_size_t),
                ("lookup_count", ctypes.c_size_t),
                ("lookup_input_graphs", ctypes.c_void_p), ("lookup_table_graphs", ctypes.c_void_p),
                ("lookup_graphs", ctypes.c_void_p), ("trash_count", ctypes.c_size_t),
                ("trash_graphs", ctypes.c_void_p), ("trash_selectors", ctypes.c_void_p)]


class PlonkKey:
    """Proving key of a PLONK circuit (custom gates, copy permutation and
    optionally lookups, trashcans and instance columns), resident on the
    device (midnight_hip_plonk_keygen_with_arguments); proofs with prove().

    lookups: list of (input compression graph, table compression graph,
    numerator lookup graph); trash: list of (graph, selector fixed column)."""

    def __init__(self, k, srs, degree, fixed, advice_cols, advice_queries, fixed_queries, perm_columns,
                 perm_mapping, gate, description=b"", instance_cols=0, lookups=(), trash=(), committed_instances=0,
                 instance_queries=()):
        n = 1 << k
        fx = _fr_array(fixed) if fixed is not None and len(fixed) else None
        aq = np.ascontiguousarray([(c, r & 0xFFFFFFFF) for c, r in advice_queries], dtype=np.uint32).reshape(-1)
        fq = np.ascontiguousarray([(c, r & 0xFFFFFFFF) for c, r in fixed_queries], dtype=np.uint32).reshape(-1)
        col_arr, col_ptr = _column_refs(perm_columns)
        mp = np.ascontiguousarray(perm_mapping, dtype=np.uint32)
        assert mp.size == 2 * len(perm_columns) * n
        h = ctypes.c_void_p()
        lookups, trash = list(lookups), list(trash)
        li_arr, li_ptr = _graphs([lk[0] for lk in lookups])
        lt_arr, lt_ptr = _graphs([lk[1] for lk in lookups])
        lg_arr, lg_ptr = _graphs([lk[2] for lk in lookups])
        tg_arr, tg_ptr = _graphs([t[0] for t in trash])
        tsel = np.ascontiguousarray([t[1] for t in trash], dtype=np.uint32)
        iq = np.ascontiguousarray([(c, r & 0xFFFFFFFF) for c, r in instance_queries], dtype=np.uint32).reshape(-1)
        args = RawPlonkArguments(instance_cols, committed_instances, iq.ctypes.data if len(iq) else None,
                                 len(instance_queries), len(lookups), li_ptr, lt_ptr, lg_ptr, len(trash), tg_ptr,
                                 tsel.ctypes.data if len(trash) else None)
        check(lib().midnight_hip_plonk_keygen_with_arguments(
            ctypes.byref(h), k, srs.handle, srs.npoints, degree, _ptr(fx), 0 if fx is None else fx.shape[0] // n,
            advice_cols, _ptr(aq), len(advice_queries), _ptr(fq), len(fixed_queries), col_ptr, len(perm_columns),
            mp.ctypes.data, gate.as_raw(), description, ctypes.byref(args)))
        self.instance_cols = instance_cols
        self.handle = h
        self.k, self.n = k, n
        self.advice_cols = advice_cols
        self.nfixed = 0 if fx is None else fx.shape[0] // n
        self.nperm = len(perm_columns)

    def vk(self):
        fc = np.zeros((self.nfixed, 18), dtype=np.uint64)
        pc = np.zeros((self.nperm, 18), dtype=np.uint64)
        rp = np.zeros(4, dtype=np.uint64)
        check(lib().midnight_hip_plonk_vk(self.handle, P(fc) if self.nfixed else None, P(pc), P(rp)))
        return fc, pc, rp

    def set_graphs(self, enable):
        check(lib().midnight_hip_plonk_set_graphs(self.handle, enable))

    def set_rng_threads(self, threads):
        check(lib().midnight_hip_plonk_set_rng_threads(self.handle, threads))

    def proof_size(self):
        ln = c_size()
        check(lib().midnight_hip_plonk_proof_size(self.handle, ctypes.byref(ln)))
        return ln.value

    def prove(self, advice, seed, capacity=None, instance=None):
        """instance: list of instance_cols arrays of Montgomery values (or
        None: every column empty)"""
        a = _fr_array(advice)
        if capacity is None:
            capacity = self.proof_size()
        buf = ctypes.create_string_buffer(max(capacity, 1))
        ln = c_size()
        if instance is None and self.instance_cols == 0:
            check(lib().midnight_hip_plonk_prove(self.handle, P(a), a.shape[0], seed, buf, capacity,
                                                 ctypes.byref(ln)))
            return buf.raw[:ln.value]
        cols = [np.ascontiguousarray(c, dtype=np.uint64).reshape(-1, 4) for c in (instance or [])]
        lens = np.ascontiguousarray([c.shape[0] for c in cols] or [0] * self.instance_cols, dtype=np.uint64)
        vals = np.concatenate(cols) if cols and sum(c.shape[0] for c in cols) else None
        check(lib().midnight_hip_plonk_prove_with_instances(
            self.handle, P(a), a.shape[0], None if vals is None else vals.ctypes.data,
            lens.ctypes.data if len(lens) else None, len(lens), seed, buf, capacity, ctypes.byref(ln)))
        return buf.raw[:ln.value]

    def free(self):
        if self.handle and self.handle.value:
            lib().midnight_hip_plonk_key_free(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ProofServer:
    """Continuous-batching proof server (midnight_hip_proof_server_*)."""

    def __init__(self, slots, hbm_budget_bytes):
        h = ctypes.c_void_p()
        check(lib().midnight_hip_proof_server_create(ctypes.byref(h), slots, hbm_budget_bytes))
        self.handle = h

    def submit(self, pk, advice, seed, instance=None):
        """instance: as PlonkKey.prove takes it (None: every column empty)"""
        a = _fr_array(advice)
        jid = ctypes.c_uint64()
        if instance is None:
            check(lib().midnight_hip_proof_server_submit(self.handle, pk.handle, P(a), a.shape[0], seed,
                                                         ctypes.byref(jid)))
            return jid.value
        cols = [np.ascontiguousarray(c, dtype=np.uint64).reshape(-1, 4) for c in instance]
        lens = np.ascontiguousarray([c.shape[0] for c in cols], dtype=np.uint64)
        vals = np.concatenate(cols) if sum(c.shape[0] for c in cols) else None
        check(lib().midnight_hip_proof_server_submit_with_instances(
            self.handle, pk.handle, P(a), a.shape[0], None if vals is None else vals.ctypes.data,
            lens.ctypes.data if len(lens) else None, len(lens), seed, ctypes.byref(jid)))
        return jid.value

    def wait(self, job, capacity=1 << 16):
        buf = ctypes.create_string_buffer(max(capacity, 1))
        ln = c_size()
        qms, pms = ctypes.c_double(), ctypes.c_double()
        check(lib().midnight_hip_proof_server_wait(self.handle, job, buf, capacity, ctypes.byref(ln),
                                                   ctypes.byref(qms), ctypes.byref(pms)))
        return buf.raw[:ln.value], qms.value, pms.value

    def free(self):
        if self.handle and self.handle.value:
            lib().midnight_hip_proof_server_free(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
