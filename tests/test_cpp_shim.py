"""The C++ drop-in header (include/vecsz_interface.hh) compiles against the
C ABI and runs the reference's call sequence: without a GPU it reports that
loudly; on a GPU (marked) compress -> decompress round-trips within the
error bound, with and without the lossless pass."""
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


@pytest.mark.gpu
@pytest.mark.parametrize("lossless", ["none", "gzip", "zstd"])
def test_shim_round_trip_on_gpu(tmp_path, lossless):
    exe = _build(tmp_path)
    out = subprocess.run([exe, lossless], capture_output=True, text=True, timeout=120).stdout.split()
    assert out[0] == "ok", out
    size, lsize, maxerr, eb = int(out[1]), int(out[2]), float(out[3]), float(out[4])
    assert size > 0 and (lsize > 0) == (lossless != "none")
    # the reference's f32 reciprocal and rounding allow |x| * 2^-22 beyond eb (D15); |x| <= 3.01 here
    assert maxerr <= eb + 3.01 * 2.0 ** -22
    assert out[5:] == ["3", "40", "20"]  # Decompress set ndim / dims from the archive


@pytest.mark.gpu
@pytest.mark.parametrize("padding", ["block,0", "edge,2", "global,3"])
def test_shim_round_trip_padded_on_gpu(tmp_path, padding):
    """ap->szwf.*_padding / pad_type / pad_constant reach the library."""
    exe = _build(tmp_path)
    out = subprocess.run([exe, "none", padding], capture_output=True, text=True, timeout=120).stdout.split()
    assert out[0] == "ok", out
    maxerr, eb = float(out[3]), float(out[4])
    assert maxerr <= eb + 3.01 * 2.0 ** -22
