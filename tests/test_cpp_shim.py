"""The C++ drop-in header (include/vecsz_interface.hh) compiles against the
C ABI and runs the reference's call sequence: without a GPU it reports that
loudly; on a GPU (marked) Compress writes the oracle's archive byte for byte
(after undoing the lossless pass, if any), Decompress returns the oracle's
floats bit for bit and fills ap's ndim, dims, len, nblk4, stride4 and
padding flags as the reference's Decompress and InitializeDims do."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "vecsz_amd")


def _build(tmp_path):
    exe = str(tmp_path / "shim_roundtrip")
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "shim_roundtrip.cc"), "-L", LIBDIR, "-lvecsz_hip",
           f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_shim_compiles_and_fails_loudly_without_gpu(tmp_path):
    exe = _build(tmp_path)
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the round trip runs in the gpu test")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120).stdout.strip()
    assert out == "nodevice"


def _run_on_field(tmp_path, lossless, padding):
    import numpy as np

    from oracle import pyoracle as po
    from vecsz_amd import lossless_decompress

    exe = _build(tmp_path)
    x = np.random.default_rng(11).standard_normal((20, 30, 40)).cumsum(axis=2).astype(np.float32)
    x.tofile(tmp_path / "in.f32")
    print(f"shim round trip: lossless {lossless} padding {padding}", flush=True)
    out = subprocess.run([exe, lossless, padding, str(tmp_path)], capture_output=True, text=True,
                         timeout=120).stdout.split()
    assert out[0] == "ok", out
    size, lsize = int(out[1]), int(out[2])
    assert (lsize > 0) == (lossless != "none")
    body = (tmp_path / "out.bin").read_bytes()
    payload = lossless_decompress(body, size) if lsize else body
    if padding == "none":
        want, _, _ = po.compress(x, 1e-3, False)
    else:
        mode, t = padding.split(",")
        want, _, _ = po.compress(x, 1e-3, False, pad_mode=mode, pad_type=["max", "min", "avg", "const"][int(t)],
                                 pad_constant=0.75)
    assert payload == want
    y = np.fromfile(tmp_path / "out.f32", dtype=np.float32)
    y_ref = po.decompress(want, x.shape, False).ravel()
    assert np.array_equal(y.view(np.uint32), y_ref.view(np.uint32))
    # ndim, dims, then len / nblk4._0 / stride4._2 as InitializeDims sets them
    # ((d + B + 1) / B blocks), then the padding flags from the pad count
    flags = {"none": 0, "global": 1, "block": 2, "edge": 4}[padding.split(",")[0]]
    assert out[5:] == ["3", "40", "20", str(40 * 30 * 20), str((40 + 9) // 8), str(40 * 30), str(flags)]


@pytest.mark.gpu
@pytest.mark.parametrize("lossless", ["none", "gzip", "zstd"])
def test_shim_round_trip_on_gpu(tmp_path, lossless):
    _run_on_field(tmp_path, lossless, "none")


@pytest.mark.gpu
@pytest.mark.parametrize("padding", ["block,0", "edge,2", "global,3"])
def test_shim_round_trip_padded_on_gpu(tmp_path, padding):
    """ap->szwf.*_padding / pad_type / pad_constant reach the library."""
    _run_on_field(tmp_path, "none", padding)
