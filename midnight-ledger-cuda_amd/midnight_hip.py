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
        "midnight_hip_field_mul_rate": (
            E, [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "midnight_hip_fr_device_vec_to_montgomery": (E, [vp]),
        "midnight_hip_g1_msm_context_create_from_multiples": (E, [vpp, vp]),
        "midnight_hip_g1_msm_context_points_download": (E, [u64p, vp, c_size]),
        "midnight_hip_g1_msm_context_last_stats": (
            E, [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32)]),
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

    def __init__(self, points, config=None, _handle=None, _npoints=0):
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
        if self.handle and self.handle.value:
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


def field_mul_rate(blocks=0):
    fp = ctypes.c_double()
    fr = ctypes.c_double()
    check(lib().midnight_hip_field_mul_rate(blocks, ctypes.byref(fp), ctypes.byref(fr)))
    return fp.value, fr.value
