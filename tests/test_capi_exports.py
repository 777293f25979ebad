"""The C-ABI library loads and exports every function include/midnight_hip.h
declares (no GPU needed: nothing is computed).  Also checks that the
argument errors the reference returns before touching the device come back
as nonzero RustError codes with a message."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "midnight_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(midnight_(?:sppark|hip)_\w+)\s*\(", text)))


def test_header_declares_the_hot_path():
    names = declared_functions()
    for required in ("midnight_sppark_bls12_381_g1_msm", "midnight_sppark_bls12_381_fr_ntt",
                     "midnight_sppark_bls12_381_g1_msm_context_device_scrambled_batch",
                     "midnight_sppark_bls12_381_fr_divide_ntt_zeta"):
        assert required in names


def test_library_exports_every_declared_symbol(H):
    lib = H.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_argument_errors_without_device(H):
    L = H.lib()
    h = ctypes.c_void_p()
    # null out pointer (bls12_381_g1_msm.cu:2748-2751)
    err = L.midnight_sppark_bls12_381_fr_device_vec_upload(None, None, 0)
    assert err.code != 0 and ctypes.string_at(err.message) == b"out is null"
    # invalid lane (bls12_381_g1_msm.cu:136-137)
    with pytest.raises(H.MidnightHipError, match="invalid sppark proof execution lane"):
        H.set_lane(99)
    # null MSM context (bls12_381_g1_msm.cu:6295-6297)
    with pytest.raises(H.MidnightHipError, match="null BLS12-381 G1 MSM context"):
        H.check(L.midnight_sppark_bls12_381_g1_msm_context_batch(None, None, None, 1, 1))
    del h


def test_numerator_argument_errors_without_device(H):
    import numpy as np
    z = np.zeros(4, dtype=np.uint64)
    # run_custom = 0 needs initial_out (bls12_381_g1_msm.cu:5245-5247)
    with pytest.raises(H.MidnightHipError, match="initial_out"):
        H.eval_numerator(16, run_custom=False, beta=z, gamma=z, theta=z, trash_challenge=z, y=z, omega=z,
                         delta_start=z, delta_const=z, rot_scale=1, last_rotation=-1, chunk_len=3)
    L = H.lib()
    err = L.midnight_sppark_bls12_381_numerator_context_eval(
        None, None, None, 0, 0, 0, 0, 0, None, 0, *([H.RawFr()] * 5), *([0] * 8))
    assert err.code != 0
    H.check(H.RustError(0, None))
