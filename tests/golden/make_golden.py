"""Writes tests/golden/golden.npz + golden.json from the reference itself.

Runs the reference's own Compress/Decompress (compiled from its sources into
oracle/_ref/libvecsz_ref.so by oracle/Makefile.ref) on small seeded inputs
chosen to exercise every per-element rule of the path, and stores inputs,
archives and decompressed fields.  1D decompression uses the reference's
pieces with the compressor's +0.0 pad (the reference's own 1D Decompress reads
an unallocated pad value).  For cases whose archives are too large to commit
(full-size fields, a tree of more than 65536 nodes) only SHA-256 digests of
the integer-exact input and of the reference archive are stored.

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402

from tests.golden.cases import digest_cases, pad_cases, small_cases  # noqa: E402


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def main():
    if not po.ref_available():
        raise SystemExit("oracle/_ref/libvecsz_ref.so missing: make -f oracle/Makefile.ref")
    arrays = {}
    manifest = {"cases": [], "digests": []}
    for c in small_cases():
        x = c["x"]
        a, nout = po.ref_compress(x, c["eb"], c["r2r"], dict_size=c["dict"], block=c["block"])
        y = po.ref_decompress(a, x, c["r2r"], dict_size=c["dict"], block=c["block"], pad0=(x.ndim == 1))
        name = c["name"]
        arrays[name + "__x"] = x
        arrays[name + "__archive"] = np.frombuffer(a, dtype=np.uint8)
        arrays[name + "__out"] = y.reshape(x.shape)
        manifest["cases"].append({
            "name": name, "shape": list(x.shape), "eb": c["eb"], "r2r": c["r2r"], "dict": c["dict"],
            "block": c["block"], "n_outlier": int(nout), "archive_bytes": len(a),
            "value_range": po.value_range(x) if c["r2r"] else 0.0, "sha256_archive": sha(a),
        })
        print(f"{name:32s} {str(x.shape):16s} archive={len(a)} n_out={nout}")
    # The reference appends pad values through one counter shared by its
    # OpenMP threads: one thread gives the serial block order.
    threads = po.ref().ref_set_threads(0)
    po.ref().ref_set_threads(1)
    for c in pad_cases():
        x = c["x"]
        kw = dict(pad_mode=c["pad_mode"], pad_type=c["pad_type"], pad_constant=c["pad_constant"])
        a, nout = po.ref_compress(x, c["eb"], c["r2r"], dict_size=c["dict"], block=c["block"], threads=1, **kw)
        y = po.ref_decompress(a, x, c["r2r"], dict_size=c["dict"], block=c["block"])
        name = c["name"]
        arrays[name + "__x"] = x
        arrays[name + "__archive"] = np.frombuffer(a, dtype=np.uint8)
        arrays[name + "__out"] = y.reshape(x.shape)
        manifest["cases"].append({
            "name": name, "shape": list(x.shape), "eb": c["eb"], "r2r": c["r2r"], "dict": c["dict"],
            "block": c["block"], "n_outlier": int(nout), "archive_bytes": len(a),
            "value_range": po.value_range(x) if c["r2r"] else 0.0, "sha256_archive": sha(a), **kw,
        })
        print(f"{name:32s} {str(x.shape):16s} archive={len(a)} n_out={nout}")
    po.ref().ref_set_threads(threads)
    for c in digest_cases():
        x = c["x"]
        a, nout = po.ref_compress(x, c["eb"], c["r2r"], dict_size=c["dict"], block=c["block"])
        y = po.ref_decompress(a, x, c["r2r"], dict_size=c["dict"], block=c["block"], pad0=(x.ndim == 1))
        manifest["digests"].append({
            "name": c["name"], "shape": list(x.shape), "eb": c["eb"], "r2r": c["r2r"], "dict": c["dict"],
            "block": c["block"], "n_outlier": int(nout), "archive_bytes": len(a),
            "sha256_input": sha(x.tobytes()), "sha256_archive": sha(a), "sha256_out": sha(y.tobytes()),
            "node_count": int.from_bytes(po.archive_sections(a)["huff"][:4], "big"),
        })
        print(f"{c['name']:32s} {str(x.shape):16s} archive={len(a)} n_out={nout} (digest only)")
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **arrays)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
