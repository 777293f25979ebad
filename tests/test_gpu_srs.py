"""GPU: KZG parameter files (bls_midnight_2pK, poly/kzg/params.rs:170-262)
parsed on the device into resident MSM contexts (midnight_hip_srs_*).

The file is rebuilt byte for byte from the reference's 2^14 SRS data
(tests/golden/srs_2p14.g1c holds its G1 points, srs_2p14_g2.bin its G2
tail; test_oracle_g1_rng.py checks the rebuild against the original file
where the reference tree is present).  Checked: every point equals the
oracle's parse; a commitment of Lagrange values against g_lagrange
(commit_lagrange, poly/kzg/mod.rs:1464-1536) equals the commitment of their
inverse NTT against g and the oracle's; the three serde formats give the
same points; invalid encodings (off the curve, outside the subgroup,
coordinates not below p, wrong flags, truncated files) are rejected where
the reference's reader rejects them.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
P_MOD = 0x1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab


@pytest.fixture(scope="module")
def fixture(O):
    k, g, gl = O.load_srs_fixture(os.path.join(GOLDEN, "srs_2p14.g1c"))
    with open(os.path.join(GOLDEN, "srs_2p14_g2.bin"), "rb") as f:
        tail = f.read()
    raw = k.to_bytes(4, "little") + b"".join(O.to_uncompressed(p) for p in g) + \
        b"".join(O.to_uncompressed(p) for p in gl) + tail
    with open(os.path.join(GOLDEN, "srs_2p14.g1c"), "rb") as f:
        comp = f.read()
    processed = comp + tail[:192]      # compressed G1; the G2 bytes are carried, not parsed
    return k, g, gl, tail, raw, processed


def _point_off(k, i, lagrange=False, size=96):
    return 4 + ((1 << k) if lagrange else 0) * size + i * size


def test_srs_raw_bytes_points_and_commit_lagrange(H, O, fixture):
    k, g, gl, tail, raw, _ = fixture
    n = 1 << k
    srs = H.Srs(raw, format=H.Srs.RAW_BYTES)
    assert srs.k == k
    assert np.array_equal(srs.g.points(n), g)
    assert np.array_equal(srs.g_lagrange.points(n), gl)
    assert srs.g2() == tail
    v = O.random_fr(n, 1414)
    coeffs = H.fr_ntt(v, k, inverse=True)
    dv = H.DeviceVec.upload(np.concatenate([v, coeffs]))
    lag = srs.g_lagrange.device_batch(dv, 0, n, 1)[0]
    mono = srs.g.device_batch(dv, n, n, 1)[0]
    assert O.g1_eq(lag, mono)
    assert O.g1_eq(lag, O.msm(g, O.ifft(v, k)))
    # a prefix of the monomial SRS (smaller circuits commit against it)
    part = srs.g.device_batch(dv, n, 1 << 10, 1)[0]
    assert O.g1_eq(part, O.msm(g[:1 << 10], coeffs[:1 << 10]))
    dv.free()
    srs.free()


def test_srs_formats_agree(H, O, fixture, tmp_path):
    k, g, gl, tail, raw, processed = fixture
    n = 1 << k
    for data, fmt in ((raw, H.Srs.RAW_BYTES_UNCHECKED), (processed, H.Srs.PROCESSED)):
        srs = H.Srs(data, format=fmt)
        assert np.array_equal(srs.g.points(n), g), fmt
        assert np.array_equal(srs.g_lagrange.points(n), gl), fmt
        srs.free()
    path = tmp_path / "bls_midnight_2p14"
    path.write_bytes(raw)
    srs = H.Srs(path=str(path), format=H.Srs.RAW_BYTES)
    assert np.array_equal(srs.g_lagrange.points(n), gl)
    srs.free()
    with pytest.raises(H.MidnightHipError):
        H.Srs(path=str(tmp_path / "missing"), format=H.Srs.RAW_BYTES)


def _off_subgroup_point(O):
    """an affine point on y^2 = x^3 + 4 outside the r-torsion subgroup
    (a random curve point is in it with probability 1 / cofactor)"""
    for x in range(1, 200):
        xm = np.zeros(6, dtype=np.uint64)
        if not O.lib().orc_fp_from_canonical(O.P(xm), O.P(np.array([x, 0, 0, 0, 0, 0], dtype=np.uint64))):
            continue
        rhs = O.fp_op("add", O.fp_op("mul", O.fp_op("mul", xm, xm), xm), _fp_int(O, 4))
        y, ok = O.fp_sqrt(rhs)
        if ok:
            p = np.concatenate([xm, y])
            # outside the subgroup: [r - 1]P is not -P
            minus_one = O.fr_op("neg", O.fr_from_int(1))
            assert not O.g1_eq(O.g1_mul(p, minus_one), _neg(O, p))
            return p
    raise AssertionError("no curve point found")


def _fp_int(O, v):
    out = np.zeros(6, dtype=np.uint64)
    O.lib().orc_fp_from_canonical(O.P(out), O.P(np.array([v, 0, 0, 0, 0, 0], dtype=np.uint64)))
    return out


def test_srs_rejects_invalid_points(H, O, fixture):
    k, g, gl, tail, raw, processed = fixture
    n = 1 << k

    def patched(base, off, new):
        b = bytearray(base)
        b[off:off + len(new)] = new
        return bytes(b)

    # off the curve: one low bit of x of g[5]
    off = _point_off(k, 5) + 47
    bad = patched(raw, off, bytes([raw[off] ^ 1]))
    with pytest.raises(H.MidnightHipError, match=r"g\[5\]"):
        H.Srs(bad, format=H.Srs.RAW_BYTES)
    srs = H.Srs(bad, format=H.Srs.RAW_BYTES_UNCHECKED)     # the unchecked reader takes it as it is
    srs.free()
    # on the curve, outside the subgroup: g_lagrange[7]
    q = _off_subgroup_point(O)
    bad = patched(raw, _point_off(k, 7, lagrange=True), O.to_uncompressed(q))
    with pytest.raises(H.MidnightHipError, match=r"g_lagrange\[7\]"):
        H.Srs(bad, format=H.Srs.RAW_BYTES)
    # a coordinate not below p
    bad = patched(raw, _point_off(k, 3), P_MOD.to_bytes(48, "big"))
    for fmt in (H.Srs.RAW_BYTES, H.Srs.RAW_BYTES_UNCHECKED):
        with pytest.raises(H.MidnightHipError):
            H.Srs(bad, format=fmt)
    # the compression flag in an uncompressed file; a sign flag there
    for flag in (0x80, 0x20):
        off = _point_off(k, 9)
        bad = patched(raw, off, bytes([raw[off] | flag]))
        with pytest.raises(H.MidnightHipError):
            H.Srs(bad, format=H.Srs.RAW_BYTES_UNCHECKED)
    # the identity encoding is a valid point (all zero in the context)
    ident = bytes([0x40]) + bytes(95)
    srs = H.Srs(patched(raw, _point_off(k, 11), ident), format=H.Srs.RAW_BYTES)
    pts = srs.g.points(16)
    assert not pts[11].any() and np.array_equal(pts[12], g[12])
    srs.free()
    # compressed: the sign bit picks -P; a missing compression flag is rejected
    off = _point_off(k, 2, size=48)
    flipped = patched(processed, off, bytes([processed[off] ^ 0x20]))
    srs = H.Srs(flipped, format=H.Srs.PROCESSED)
    got = srs.g.points(4)
    assert np.array_equal(got[2], O.to_affine(_neg(O, g[2])))
    srs.free()
    with pytest.raises(H.MidnightHipError):
        H.Srs(patched(processed, off, bytes([processed[off] & 0x7f])), format=H.Srs.PROCESSED)
    # truncated file, unknown format
    with pytest.raises(H.MidnightHipError):
        H.Srs(raw[:-1], format=H.Srs.RAW_BYTES)
    with pytest.raises(H.MidnightHipError):
        H.Srs(raw, format=3)


def _neg(O, p_aff):
    j = O.g1j()
    O.lib().orc_g1j_from_affine(O.P(j), O.P(np.ascontiguousarray(p_aff)))
    out = O.g1j()
    O.lib().orc_g1j_neg(O.P(out), O.P(j))
    return out
