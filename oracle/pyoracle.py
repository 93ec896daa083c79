"""TEST INFRASTRUCTURE ONLY: ctypes bindings for the C oracle (liboracle.so)
and, where it was built, the reference library (oracle/_ref/libvecsz_ref.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORC_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libvecsz_ref.so")

_c_size_p = ctypes.POINTER(ctypes.c_size_t)
_c_vp = ctypes.c_void_p


class OracleError(RuntimeError):
    pass


def _load_oracle():
    if not os.path.exists(_ORC_PATH):
        raise OracleError(f"{_ORC_PATH} missing: run `make -C oracle`")
    lib = ctypes.CDLL(_ORC_PATH)
    lib.orc_value_range.argtypes = [_c_vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)]
    lib.orc_compress.argtypes = [ctypes.c_int, _c_size_p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_int, _c_vp, ctypes.POINTER(_c_vp), _c_size_p, _c_size_p,
                                 ctypes.POINTER(ctypes.c_double)]
    lib.orc_decompress.argtypes = [_c_vp, ctypes.c_size_t, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, _c_vp, ctypes.c_size_t]
    lib.orc_dualquant.argtypes = [ctypes.c_int, _c_size_p, ctypes.c_int, ctypes.c_int, _c_vp, _c_vp, _c_vp, _c_vp]
    lib.orc_huffman_codebook.argtypes = [_c_vp, ctypes.c_int, _c_vp, _c_vp, ctypes.POINTER(ctypes.c_int),
                                         _c_vp, _c_size_p]
    lib.orc_huffman_decode.argtypes = [_c_vp, ctypes.c_size_t, ctypes.c_size_t, _c_vp]
    lib.orc_huffman_pack.argtypes = [_c_vp, ctypes.c_size_t, _c_vp, _c_vp, _c_vp]
    lib.orc_huffman_pack.restype = ctypes.c_size_t
    lib.orc_reconstruct.argtypes = [ctypes.c_int, _c_size_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    _c_vp, _c_vp, _c_vp]
    lib.orc_eb_family.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _c_vp]
    lib.orc_compress_pad.argtypes = [ctypes.c_int, _c_size_p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _c_vp,
                                     ctypes.POINTER(_c_vp), _c_size_p, _c_size_p, ctypes.POINTER(ctypes.c_double)]
    lib.orc_pad_values.argtypes = [ctypes.c_int, _c_size_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float, _c_vp, _c_vp, _c_vp, _c_size_p, _c_vp]
    lib.orc_pad_halo_from_archive.argtypes = [ctypes.c_int, _c_size_p, ctypes.c_int, _c_vp, ctypes.c_uint64, _c_vp]
    lib.orc_dualquant_halo.argtypes = [ctypes.c_int, _c_size_p, ctypes.c_int, ctypes.c_int, _c_vp, _c_vp, _c_vp,
                                       _c_vp, _c_vp]
    lib.orc_nblk.argtypes = [ctypes.c_size_t, ctypes.c_int]
    lib.orc_nblk.restype = ctypes.c_size_t
    lib.orc_free.argtypes = [_c_vp]
    return lib


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        _orc = _load_oracle()
    return _orc


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise OracleError(f"{_REF_PATH} missing: run `make -f oracle/Makefile.ref`")
        lib = ctypes.CDLL(_REF_PATH)
        lib.ref_compress.argtypes = [ctypes.c_int, _c_size_p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_char_p, _c_vp, ctypes.POINTER(_c_vp), _c_size_p,
                                     _c_size_p, ctypes.c_int]
        lib.ref_decompress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, _c_vp,
                                       ctypes.c_size_t, ctypes.POINTER(_c_vp), _c_size_p, ctypes.c_int]
        lib.ref_decompress_pad0.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, _c_vp,
                                            ctypes.POINTER(_c_vp), _c_size_p, ctypes.c_int]
        lib.ref_compress_pad.argtypes = [ctypes.c_int, _c_size_p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_char_p,
                                         _c_vp, ctypes.POINTER(_c_vp), _c_size_p, _c_size_p, ctypes.c_int]
        lib.ref_free.argtypes = [_c_vp]
        lib.ref_free_f32.argtypes = [_c_vp]
        lib.ref_set_threads.argtypes = [ctypes.c_int]
        lib.ref_set_threads.restype = ctypes.c_int
        lib.ref_lossless_compress.argtypes = [ctypes.c_int, _c_vp, ctypes.c_ulong, ctypes.POINTER(_c_vp)]
        lib.ref_lossless_compress.restype = ctypes.c_ulong
        lib.ref_lossless_decompress.argtypes = [_c_vp, ctypes.c_ulong, ctypes.POINTER(_c_vp), ctypes.c_ulong]
        lib.ref_lossless_decompress.restype = ctypes.c_long
        lib.ref_free_malloc.argtypes = [_c_vp]
        _ref = lib
    return _ref


class EbFamily(ctypes.Structure):
    _fields_ = [("eb_final", ctypes.c_double), ("r_d", ctypes.c_double), ("twoeb_d", ctypes.c_double),
                ("r_f", ctypes.c_float)]


def _dims_arr(shape_xfirst):
    d = [int(v) for v in shape_xfirst] + [1] * (3 - len(shape_xfirst))
    return (ctypes.c_size_t * 3)(*d)


def dims_xfirst(arr: np.ndarray):
    """numpy C-order shape (slowest first) -> reference dims (x fastest first)."""
    return tuple(int(v) for v in reversed(arr.shape))


def value_range(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = ctypes.c_double()
    rc = orc().orc_value_range(x.ctypes.data, x.size, ctypes.byref(out))
    if rc:
        raise OracleError(f"orc_value_range rc={rc}")
    return out.value


def eb_family(r2r: bool, eb: float, rng: float = 0.0) -> EbFamily:
    f = EbFamily()
    rc = orc().orc_eb_family(int(r2r), eb, rng, ctypes.byref(f))
    if rc:
        raise OracleError(f"orc_eb_family rc={rc}")
    return f


PAD_MODES = {"none": 0, "global": 1, "block": 2, "edge": 3}
PAD_TYPES = {"max": 0, "min": 1, "avg": 2, "const": 3}


def compress(x: np.ndarray, eb: float, r2r: bool = False, dict_size: int = 4096, block: int = 8,
             pad_mode: str = "none", pad_type: str = "const", pad_constant: float = 0.0):
    """Oracle compress. Returns (archive bytes, n_outlier, eb_final)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    dims = dims_xfirst(x)
    out = _c_vp()
    size = ctypes.c_size_t()
    nout = ctypes.c_size_t()
    ebf = ctypes.c_double()
    rc = orc().orc_compress_pad(len(dims), _dims_arr(dims), int(r2r), eb, dict_size, block, PAD_MODES[pad_mode],
                                PAD_TYPES[pad_type], pad_constant, x.ctypes.data, ctypes.byref(out),
                                ctypes.byref(size), ctypes.byref(nout), ctypes.byref(ebf))
    if rc:
        raise OracleError(f"orc_compress rc={rc}")
    data = ctypes.string_at(out.value, size.value)
    orc().orc_free(out)
    return data, nout.value, ebf.value


def decompress(archive: bytes, shape, r2r: bool = False, value_rng: float = 0.0, dict_size: int = 4096,
               block: int = 8) -> np.ndarray:
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.float32)
    buf = np.frombuffer(archive, dtype=np.uint8)
    rc = orc().orc_decompress(buf.ctypes.data, len(archive), int(r2r), value_rng, dict_size, block,
                              out.ctypes.data, n)
    if rc:
        raise OracleError(f"orc_decompress rc={rc}")
    return out.reshape(shape)


def dualquant(x: np.ndarray, eb: float, r2r: bool = False, dict_size: int = 4096, block: int = 8):
    """Oracle dual quantization: returns (int32 codes, float32 outlier scratch)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    dims = dims_xfirst(x)
    rng = value_range(x) if r2r else 0.0
    f = eb_family(r2r, eb, rng)
    code = np.empty(x.shape, dtype=np.int32)
    outl = np.empty(x.shape, dtype=np.float32)
    rc = orc().orc_dualquant(len(dims), _dims_arr(dims), block, dict_size // 2, ctypes.byref(f),
                             x.ctypes.data, code.ctypes.data, outl.ctypes.data)
    if rc:
        raise OracleError(f"orc_dualquant rc={rc}")
    return code, outl


def nblk(d: int, block: int) -> int:
    """Blocks along a dimension as the reference counts them: (d + B + 1) // B."""
    return int(orc().orc_nblk(d, block))


def pad_values(x: np.ndarray, eb: float, r2r: bool = False, block: int = 8, pad_mode: str = "block",
               pad_type: str = "max", pad_constant: float = 0.0):
    """Oracle pad values: (archive pad_vals float32[pad_len], halo float32[nblocks, 3] (z, y, x faces))."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    dims = dims_xfirst(x)
    f = eb_family(r2r, eb, value_range(x) if r2r else 0.0)
    plen = ctypes.c_size_t()
    args = (len(dims), _dims_arr(dims), block, PAD_MODES[pad_mode], PAD_TYPES[pad_type], pad_constant,
            ctypes.byref(f), x.ctypes.data)
    rc = orc().orc_pad_values(*args, None, ctypes.byref(plen), None)
    if rc:
        raise OracleError(f"orc_pad_values rc={rc}")
    nb = int(np.prod([nblk(d, block) for d in dims]))
    pv = np.zeros(max(plen.value, 1), dtype=np.float32)
    halo = np.zeros((nb, 3), dtype=np.float32)
    rc = orc().orc_pad_values(*args, pv.ctypes.data, ctypes.byref(plen), halo.ctypes.data)
    if rc:
        raise OracleError(f"orc_pad_values rc={rc}")
    return pv[:plen.value], halo


def halo_from_archive(dims_xfirst_: tuple, block: int, pad: np.ndarray) -> np.ndarray:
    """Halo values decompress derives from an archive's pad section."""
    pad = np.ascontiguousarray(pad, dtype=np.float32)
    nb = int(np.prod([nblk(d, block) for d in dims_xfirst_]))
    halo = np.zeros((nb, 3), dtype=np.float32)
    orc().orc_pad_halo_from_archive(len(dims_xfirst_), _dims_arr(dims_xfirst_), block,
                                    pad.ctypes.data if pad.size else None, pad.size, halo.ctypes.data)
    return halo


def codebook(freq: np.ndarray, state_num: int):
    """Oracle Huffman codebook: (code_bits u64, code_len u8, node_count, tree bytes)."""
    freq = np.ascontiguousarray(freq, dtype=np.uint64)
    assert freq.size >= state_num
    cb = np.zeros(state_num, dtype=np.uint64)
    cl = np.zeros(state_num, dtype=np.uint8)
    nc = ctypes.c_int()
    tb = ctypes.c_size_t()
    rc = orc().orc_huffman_codebook(freq.ctypes.data, state_num, cb.ctypes.data, cl.ctypes.data,
                                    ctypes.byref(nc), None, ctypes.byref(tb))
    if rc:
        raise OracleError(f"orc_huffman_codebook rc={rc}")
    tree = np.zeros(tb.value, dtype=np.uint8)
    orc().orc_huffman_codebook(freq.ctypes.data, state_num, cb.ctypes.data, cl.ctypes.data, ctypes.byref(nc),
                               tree.ctypes.data, ctypes.byref(tb))
    return cb, cl, nc.value, tree.tobytes()


def huffman_decode(huff: bytes, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.int32)
    buf = np.frombuffer(huff, dtype=np.uint8)
    rc = orc().orc_huffman_decode(buf.ctypes.data, len(huff), n, out.ctypes.data)
    if rc:
        raise OracleError(f"orc_huffman_decode rc={rc}")
    return out


def archive_sections(archive: bytes):
    """Split an archive payload into its fields (datapack layout)."""
    a = np.frombuffer(archive, dtype=np.uint8)
    nd = int(a[:4].view(np.int32)[0])
    dims = tuple(int(v) for v in a[4:4 + 4 * nd].view(np.int32))
    off = 4 * (nd + 1)
    eb = float(a[off:off + 8].view(np.float64)[0])
    pad_len, huff_size, nout = (int(v) for v in a[off + 8:off + 32].view(np.uint64))
    p = off + 32
    pad = a[p:p + 4 * pad_len].view(np.float32)
    p += 4 * pad_len
    huff = archive[p:p + huff_size]
    p += huff_size
    outl = a[p:p + 4 * nout].view(np.float32)
    return dict(ndim=nd, dims=dims, eb=eb, pad_len=pad_len, huff_size=huff_size, n_outlier=nout, pad=pad,
                huff=huff, outliers=outl, meta_bytes=off + 32)


# ------------------------------------------------------------------ reference

def ref_compress(x: np.ndarray, eb: float, r2r: bool = False, dict_size: int = 4096, block: int = 8,
                 threads: int = 0, pad_mode: str = "none", pad_type: str = "const", pad_constant: float = 0.0):
    """Run the reference's own Compress<float,int> in-process (test only)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    dims = dims_xfirst(x)
    lib = ref()
    if threads:
        lib.ref_set_threads(threads)
    path = b""
    tmp = None
    if r2r:
        tmp = tempfile.NamedTemporaryFile(suffix=".f32", delete=False)
        tmp.write(x.tobytes())
        tmp.close()
        path = tmp.name.encode()
    try:
        out = _c_vp()
        size = ctypes.c_size_t()
        nout = ctypes.c_size_t()
        if pad_mode == "none":
            lib.ref_compress(len(dims), _dims_arr(dims), int(r2r), eb, dict_size, block, path, x.ctypes.data,
                             ctypes.byref(out), ctypes.byref(size), ctypes.byref(nout), 1)
        else:
            lib.ref_compress_pad(len(dims), _dims_arr(dims), int(r2r), eb, dict_size, block, PAD_MODES[pad_mode],
                                 PAD_TYPES[pad_type], pad_constant, path, x.ctypes.data, ctypes.byref(out),
                                 ctypes.byref(size), ctypes.byref(nout), 1)
        data = ctypes.string_at(out.value, size.value)
        lib.ref_free(out)
    finally:
        if tmp is not None:
            os.unlink(tmp.name)
    return data, nout.value


def ref_decompress(archive: bytes, original: np.ndarray | None = None, r2r: bool = False, dict_size: int = 4096,
                   block: int = 8, pad0: bool = False) -> np.ndarray:
    """Run the reference's Decompress<float,int> (or its pad-0 variant for 1D)."""
    lib = ref()
    path = b""
    tmp = None
    if r2r:
        assert original is not None
        tmp = tempfile.NamedTemporaryFile(suffix=".f32", delete=False)
        tmp.write(np.ascontiguousarray(original, dtype=np.float32).tobytes())
        tmp.close()
        path = tmp.name.encode()
    try:
        buf = np.frombuffer(archive + b"\0" * 64, dtype=np.uint8)
        out = _c_vp()
        n = ctypes.c_size_t()
        if pad0:
            lib.ref_decompress_pad0(int(r2r), dict_size, block, path, buf.ctypes.data, ctypes.byref(out),
                                    ctypes.byref(n), 1)
        else:
            lib.ref_decompress(int(r2r), dict_size, block, path, buf.ctypes.data, len(archive), ctypes.byref(out),
                               ctypes.byref(n), 1)
        y = np.frombuffer(ctypes.string_at(out.value, 4 * n.value), dtype=np.float32).copy()
        lib.ref_free_f32(out)
    finally:
        if tmp is not None:
            os.unlink(tmp.name)
    return y


def ref_lossless_compress(kind: str, data: bytes) -> bytes:
    """The reference's own lossless pass (src/lossless.hh:179-198, level 3)."""
    lib = ref()
    buf = np.frombuffer(data, dtype=np.uint8)
    out = _c_vp()
    n = lib.ref_lossless_compress(1 if kind == "zstd" else 0, buf.ctypes.data, len(data), ctypes.byref(out))
    res = ctypes.string_at(out.value, n)
    lib.ref_free_malloc(out)
    return res


def ref_lossless_decompress(data: bytes, target: int) -> bytes:
    """The reference's detection + inverse (src/lossless.hh:60-75, :200-218)."""
    lib = ref()
    buf = np.frombuffer(data, dtype=np.uint8)
    out = _c_vp()
    n = lib.ref_lossless_decompress(buf.ctypes.data, len(data), ctypes.byref(out), target)
    if n < 0:
        raise OracleError("not a lossless stream")
    res = ctypes.string_at(out.value, n)
    lib.ref_free_malloc(out)
    return res
