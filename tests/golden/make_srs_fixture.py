"""Writes srs_2p14.g1c: the G1 part of the reference's SRS data file
transient-crypto/static/bls_midnight_2p14, re-encoded compressed.

The source file (written by vendor/midnight-proofs-0.8.1/src/poly/kzg/
params.rs:170-185, RawBytes format) is u32 k || 2^k uncompressed G1 `g`
([tau^i]G) || 2^k uncompressed G1 `g_lagrange` || G2 || [tau]G2.  It is only
parsed as bytes here (ZCash encoding); no reference code runs.  The fixture
keeps k and both G1 vectors, 48 bytes per point, so that tests can check
MSM(g, c) == MSM(g_lagrange, NTT(c)) without the reference tree;
srs_2p14_g2.bin keeps the file's 384-byte G2 tail, so that tests can rebuild
the whole file.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402

SRC = "/root/reference/transient-crypto/static/bls_midnight_2p14"
DST = os.path.join(HERE, "srs_2p14.g1c")
DST_G2 = os.path.join(HERE, "srs_2p14_g2.bin")


def main():
    k, g, gl = oracle.load_srs_file(SRC)
    with open(DST, "wb") as f:
        f.write(k.to_bytes(4, "little"))
        for arr in (g, gl):
            for p in arr:
                f.write(oracle.compress(p))
    print("wrote", DST, os.path.getsize(DST), "bytes")
    # the G2 tail ([1]G2 || [tau]G2, uncompressed) as the file holds it, so
    # that tests can rebuild the whole file byte for byte
    with open(SRC, "rb") as f:
        data = f.read()
    tail = data[4 + 2 * (1 << k) * 96:]
    assert len(tail) == 2 * 192
    with open(DST_G2, "wb") as f:
        f.write(tail)
    print("wrote", DST_G2, len(tail), "bytes")


if __name__ == "__main__":
    main()
