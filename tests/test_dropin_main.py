"""The reference's own command-line program (src/vecsz.cc, unchanged) built
against include/vecsz_interface.hh and libvecsz_hip.so by
oracle/Makefile.dropin: its calls at src/vecsz.cc:37, 59 and 88-93 bind to
this library, not to the reference's header-only CPU path.

Without a GPU the build is checked (which header was included, which symbols
the program takes from the library, that none of the reference's
compression code is in it) and the program must fail loudly.  On a GPU
(marked) its .sz files must hold the oracle's archive byte for byte and its
decompressed output must equal the oracle's bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
OUT = os.path.join(ROOT, "oracle", "_ref", "dropin")
EXE = os.path.join(OUT, "vecsz")


def _build():
    subprocess.run(["make", "-s", "-f", "oracle/Makefile.dropin", f"REF={REF}"], cwd=ROOT, check=True,
                   capture_output=True, text=True)


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference sources not present")
def test_reference_main_binds_to_this_library():
    _build()
    includes = open(os.path.join(OUT, "vecsz.includes")).read()
    lines = [l.split()[-1] for l in includes.splitlines() if l.startswith(".")]
    assert os.path.join(ROOT, "include", "vecsz_interface.hh") in lines
    assert os.path.join(REF, "src", "vecsz_interface.hh") not in lines
    # none of the reference's compression path came in with it
    for h in ("dualquant.hh", "huffman.hh", "lossless.hh", "datapack.hh", "padding.hh"):
        assert not any(l.endswith(h) for l in lines), h
    nm = subprocess.run(["nm", "-C", EXE], check=True, capture_output=True, text=True).stdout
    undefined = {l.split()[-1] for l in nm.splitlines() if l.split()[0] == "U"}
    for sym in ("vecsz_compress_host", "vecsz_decompress_host", "vecsz_archive_info", "vecsz_compress_bound"):
        assert sym in undefined, sym
    for sym in ("HuffmanEncode", "initHuffman", "c_lorenzo", "x_lorenzo", "datapack"):
        assert sym not in nm, sym


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference sources not present")
def test_reference_main_fails_loudly_without_gpu(tmp_path):
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the gpu test runs the program")
    _build()
    x = np.random.default_rng(0).standard_normal(4000).astype(np.float32)
    x.tofile(tmp_path / "in.f32")
    r = subprocess.run([EXE, "-t", "f32", "-z", "-i", str(tmp_path / "in.f32"), "-o", str(tmp_path / "o.sz"),
                        "-l", "4000", "-m", "abs", "-e", "1e-3"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "no usable GPU" in r.stderr


def _run(args):
    r = subprocess.run([EXE, "-t", "f32", *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/dropin/vecsz not built (needs the reference sources)")
@pytest.mark.parametrize("shape,eb,lossless", [((20, 30, 40), 1e-3, None), ((1800, 3600), 1e-4, None),
                                               ((100003,), 1e-4, None), ((20, 30, 40), 1e-3, "gzip")])
def test_reference_main_on_gpu_matches_oracle(tmp_path, shape, eb, lossless):
    from vecsz_amd import datagen, lossless_decompress

    x = datagen.grf(shape, seed=7) if len(shape) == 3 else (
        datagen.sinusoids_2d(*shape, seed=7) if len(shape) == 2 else datagen.walk_spikes_1d(shape[0], seed=7))
    x = np.ascontiguousarray(x, dtype=np.float32)
    src, sz, dec = tmp_path / "in.f32", tmp_path / "in.sz", tmp_path / "out.f32"
    x.tofile(src)
    dims = ",".join(str(d) for d in reversed(x.shape))
    print(f"dropin main: shape {shape} eb {eb} lossless {lossless}", flush=True)
    _run(["-z", "-i", str(src), "-o", str(sz), "-l", dims, "-m", "abs", "-e", repr(eb)]
         + (["-L", lossless] if lossless else []))
    f = sz.read_bytes()
    lsize, dsize = np.frombuffer(f[:16], dtype=np.uint64).tolist()
    want, _, _ = po.compress(x, eb, False)
    assert dsize == len(want)
    body = f[16:16 + (lsize or dsize)]
    payload = lossless_decompress(body, dsize) if lsize else body
    assert (lsize > 0) == (lossless is not None)
    assert payload == want
    _run(["-x", "-i", str(sz), "-o", str(dec), "-m", "abs"])
    y = np.fromfile(dec, dtype=np.float32)
    y_ref = po.decompress(want, x.shape, False).ravel()
    assert y.size == x.size  # main wrote ap->len floats: Decompress set len from the archive
    assert np.array_equal(y.view(np.uint32), y_ref.view(np.uint32))
