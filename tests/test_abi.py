"""The C-ABI library without a GPU: it loads, exports every entry point the
header declares, and its host-only entry points behave."""
import ctypes
import os
import re

import numpy as np
import pytest

from vecsz_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "vecsz_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vecsz_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_version_and_strerror():
    lib = _lib.load()
    assert b"gfx950" in lib.vecsz_version()
    for s in range(8):
        assert lib.vecsz_strerror(s)


def test_params_default_match_reference_defaults():
    p = _lib.Params()
    _lib.load().vecsz_params_default(ctypes.byref(p))
    assert (p.eb, p.block, p.dict_size, p.mode) == (1e-4, 8, 4096, 0)


def test_compress_bound():
    from vecsz_amd.interface import make_params

    p = make_params((64, 64, 64), 1e-4)
    b = ctypes.c_uint64()
    assert _lib.load().vecsz_compress_bound(ctypes.byref(p), ctypes.byref(b)) == 0
    assert b.value >= 64 ** 3 * 12


def test_archive_info_on_golden(golden):
    from vecsz_amd import archive_info

    for c in golden["cases"]:
        info = archive_info(c["archive"])
        assert info["dims"] == tuple(reversed(c["x"].shape))
        assert info["eb"] == c["eb"]
        assert info["n_outlier"] == c["n_outlier"]


def test_archive_info_rejects_truncated(golden):
    from vecsz_amd import VecszError, archive_info

    a = golden["cases"][0]["archive"]
    for cut in (0, 3, 10, 30, len(a) - 1):
        with pytest.raises(VecszError):
            archive_info(a[:cut])


def test_no_gpu_context_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from vecsz_amd import Context, VecszError

    with pytest.raises(VecszError):
        Context(0)


def test_index_words():
    """Decode index size: a 6-word header, two words per group of 64
    segments of 128 symbols, then a u32 and a u16 per segment, each array
    padded to whole u64 words."""
    from vecsz_amd import make_params

    lib = _lib.load()
    w = ctypes.c_uint64()
    for dims, segs in (((1,), 1), ((128,), 1), ((129,), 2), ((512, 512, 512), 1 << 20), ((3600, 1800), 50625)):
        p = make_params(dims, 1e-4, "abs")
        assert lib.vecsz_index_words(ctypes.byref(p), ctypes.byref(w)) == 0
        groups = -(-segs // 64)
        assert w.value == 6 + 2 * groups + -(-segs // 2) + -(-segs // 4), dims
    p = make_params((8,), 1e-4, "abs")
    p.ndim = 0
    assert lib.vecsz_index_words(ctypes.byref(p), ctypes.byref(w)) != 0


def _header(ndim, dims, huff=0, nout=0, pad=0):
    """Archive metadata (datapack layout): [i32 ndim][i32 dims][f64 eb][u64 pad][u64 huff][u64 nout]."""
    return (np.array([ndim, *dims], dtype=np.int32).tobytes() + np.array([1e-4], dtype=np.float64).tobytes()
            + np.array([pad, huff, nout], dtype=np.uint64).tobytes())


def test_archive_info_rejects_wrapping_element_count():
    """Extents whose product wraps 64 bits (2^25 each gives 2^75) or exceeds
    the library's 2^37-element cap are a corrupt header, never a tiny field."""
    from vecsz_amd import VecszError, archive_info

    for dims in ((1 << 25, 1 << 25, 1 << 25), (0x7FFFFFFF, 0x7FFFFFFF, 4), ((1 << 19), (1 << 19))):
        h = _header(len(dims), dims) + bytes(64)
        with pytest.raises(VecszError):
            archive_info(h)
    # the largest accepted count parses
    h = _header(3, (1 << 13, 1 << 12, 1 << 12)) + bytes(64)
    assert archive_info(h)["dims"] == (1 << 13, 1 << 12, 1 << 12)


def test_archive_info_rejects_more_outliers_than_elements():
    from vecsz_amd import VecszError, archive_info

    h = _header(1, (4,), huff=8, nout=5) + bytes(64)
    with pytest.raises(VecszError):
        archive_info(h)


def test_sizes_reject_wrapping_element_count():
    from vecsz_amd import make_params

    lib = _lib.load()
    b = ctypes.c_uint64()
    p = make_params((1 << 25, 1 << 25, 1 << 25), 1e-4, "abs")
    assert lib.vecsz_compress_bound(ctypes.byref(p), ctypes.byref(b)) != 0
    assert lib.vecsz_index_words(ctypes.byref(p), ctypes.byref(b)) != 0
