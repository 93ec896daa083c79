"""Host lossless pass and .sz container (no GPU): gzip output is the
reference's byte for byte (same zlib 1.2.11); zstd frames decode both ways
(the system zstd is a newer release than the reference's 1.3.5, so its
compressed bytes are not compared); the container round-trips and reads the
reference's D8-style files."""
import numpy as np
import pytest

from oracle import pyoracle as po
from vecsz_amd import VecszError, lossless_compress, lossless_decompress, lossless_kind, sz_pack, sz_unpack


def _payloads(golden):
    out = [c["archive"] for c in golden["cases"]]
    rng = np.random.default_rng(0)
    out.append(rng.integers(0, 256, 300_000, dtype=np.uint8).tobytes())  # incompressible, spans 64 KiB pieces
    out.append(bytes(200_000))                                          # highly compressible
    out.append(b"x")
    return out


@pytest.mark.parametrize("kind", ["gzip", "zstd"])
def test_round_trip(golden, kind):
    for p in _payloads(golden):
        z = lossless_compress(kind, p)
        assert lossless_kind(z) == kind
        assert lossless_decompress(z, len(p)) == p


@pytest.mark.skipif(not po.ref_available(), reason="reference build (oracle/_ref) not present")
def test_gzip_matches_reference_bytes(golden):
    for p in _payloads(golden):
        assert lossless_compress("gzip", p) == po.ref_lossless_compress("gzip", p)


@pytest.mark.skipif(not po.ref_available(), reason="reference build (oracle/_ref) not present")
def test_cross_decoding_with_reference(golden):
    for p in _payloads(golden)[:12]:
        for kind in ("gzip", "zstd"):
            theirs = po.ref_lossless_compress(kind, p)
            assert lossless_decompress(theirs, len(p)) == p          # reference stream, our inverse
            ours = lossless_compress(kind, p)
            assert po.ref_lossless_decompress(ours, len(p)) == p      # our stream, reference inverse


def test_payloads_are_not_mistaken_for_lossless_streams(golden):
    for c in golden["cases"]:
        assert lossless_kind(c["archive"]) is None


def test_corrupt_stream_is_an_error():
    z = bytearray(lossless_compress("gzip", bytes(range(256)) * 400))
    z[len(z) // 2] ^= 0xFF
    with pytest.raises(VecszError):
        lossless_decompress(bytes(z), 256 * 400)


def test_sz_container(golden):
    a = golden["cases"][0]["archive"]
    f = sz_pack(a, 0, len(a))
    assert len(f) == 16 + len(a)
    assert np.frombuffer(f[:16], dtype=np.uint64).tolist() == [0, len(a)]
    assert sz_unpack(f) == (a, 0, len(a))
    z = lossless_compress("gzip", a)
    f = sz_pack(z, len(z), len(a))
    assert sz_unpack(f) == (z, len(z), len(a))
    # the reference writes data_size bytes after the header of a lossless file (D8)
    d8 = f + bytes(max(0, len(a) - len(z)))
    assert sz_unpack(d8) == (z, len(z), len(a))
    with pytest.raises(VecszError):
        sz_unpack(f[:20])


def test_sz_container_body_size_near_2_64_is_corrupt():
    """A body size read from the file near 2^64 must not wrap the bounds check."""
    for ls, ds in ((2 ** 64 - 8, 0), (0, 2 ** 64 - 15), (2 ** 64 - 1, 2 ** 64 - 1)):
        f = np.array([ls, ds], dtype=np.uint64).tobytes() + bytes(32)
        with pytest.raises(VecszError):
            sz_unpack(f)
