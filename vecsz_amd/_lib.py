"""ctypes binding of libvecsz_hip.so (the C ABI declared in include/vecsz_hip.h)."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvecsz_hip.so")

VECSZ_OK = 0
VECSZ_EINVAL = 1
VECSZ_ENOMEM = 2
VECSZ_ECORRUPT = 3
VECSZ_ECODELEN = 4
VECSZ_EHIP = 5
VECSZ_ERANGE0 = 6
VECSZ_ENODEV = 7

MODE_ABS = 0
MODE_R2R = 1

# Every symbol include/vecsz_hip.h declares.
EXPORTS = (
    "vecsz_ctx_create", "vecsz_ctx_destroy", "vecsz_strerror", "vecsz_ctx_last_error", "vecsz_params_default",
    "vecsz_compress_bound", "vecsz_compress", "vecsz_decompress", "vecsz_index_words", "vecsz_compress_indexed",
    "vecsz_decompress_indexed", "vecsz_compress_host", "vecsz_decompress_host",
    "vecsz_archive_info", "vecsz_huffman_codebook", "vecsz_compress_batch", "vecsz_decompress_batch",
    "vecsz_ctx_set_workers", "vecsz_ctx_set_batch_fused", "vecsz_autotune", "vecsz_ctx_set_launch_shape",
    "vecsz_server_create", "vecsz_server_destroy", "vecsz_server_submit_compress",
    "vecsz_server_submit_decompress", "vecsz_server_wait", "vecsz_lossless_bound", "vecsz_lossless_compress",
    "vecsz_lossless_kind", "vecsz_lossless_decompress", "vecsz_sz_pack", "vecsz_sz_unpack", "vecsz_ctx_last_sidecar",
    "vecsz_params_from_sidecar", "vecsz_verify", "vecsz_ctx_enable_timing", "vecsz_ctx_timings", "vecsz_version",
    "vecsz_compress_noalloc", "vecsz_ctx_set_fused_encode", "vecsz_ctx_set_device_codebook",
    "vecsz_ctx_set_batch_graph",
)


class Params(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("dims", ctypes.c_uint64 * 3),
        ("mode", ctypes.c_int),
        ("eb", ctypes.c_double),
        ("value_range", ctypes.c_double),
        ("block", ctypes.c_int),
        ("dict_size", ctypes.c_int),
        ("pad_mode", ctypes.c_int),
        ("pad_type", ctypes.c_int),
        ("pad_constant", ctypes.c_float),
    ]


PAD_MODES = {"none": 0, "global": 1, "block": 2, "edge": 3}
PAD_TYPES = {"max": 0, "min": 1, "avg": 2, "const": 3, "constant": 3}


VECSZ_LOSSLESS_GZIP = 1
VECSZ_LOSSLESS_ZSTD = 2


class Sidecar(ctypes.Structure):
    _fields_ = [
        ("magic", ctypes.c_uint32),
        ("version", ctypes.c_uint32),
        ("ndim", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("dict_size", ctypes.c_int32),
        ("eb", ctypes.c_double),
        ("value_range", ctypes.c_double),
        ("eb_final", ctypes.c_double),
    ]


class Stat(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("minimum", "maximum", "range", "PSNR", "MSE", "NRMSE", "coeff",
                                               "max_abserr_vs_range", "max_pwr_rel_abserr")] + [
        ("len", ctypes.c_uint64), ("max_abserr_index", ctypes.c_uint64), ("max_abserr", ctypes.c_double)]


class Tune(ctypes.Structure):
    _fields_ = [
        ("block", ctypes.c_int),
        ("dq_grid", ctypes.c_int),
        ("enc_grid", ctypes.c_int),
        ("pad_", ctypes.c_int),
        ("ms_by_block", ctypes.c_double * 4),
        ("bytes_by_block", ctypes.c_uint64 * 4),
        ("dq_ms", ctypes.c_double),
        ("enc_ms", ctypes.c_double),
    ]


class ServerResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("out_bytes", ctypes.c_uint64),
        ("n_outlier", ctypes.c_uint64),
        ("queue_ms", ctypes.c_double),
        ("run_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
    ]


class VecszError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"vecsz status {status}: {msg}")
        self.status = status


_lib = None


def load():
    """Loads the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m vecsz_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    vp, u64, u64p = ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)
    P = ctypes.POINTER(Params)
    lib.vecsz_ctx_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, vp]
    lib.vecsz_ctx_destroy.argtypes = [vp]
    lib.vecsz_strerror.argtypes = [ctypes.c_int]
    lib.vecsz_strerror.restype = ctypes.c_char_p
    lib.vecsz_ctx_last_error.argtypes = [vp]
    lib.vecsz_ctx_last_error.restype = ctypes.c_char_p
    lib.vecsz_params_default.argtypes = [P]
    lib.vecsz_params_default.restype = None
    lib.vecsz_compress_bound.argtypes = [P, u64p]
    lib.vecsz_compress.argtypes = [vp, P, vp, vp, u64, u64p, u64p]
    lib.vecsz_compress_noalloc.argtypes = [vp, P, vp, vp, u64, u64p, u64p]
    lib.vecsz_decompress.argtypes = [vp, P, vp, u64, vp, u64]
    lib.vecsz_compress_host.argtypes = [vp, P, vp, vp, u64, u64p, u64p]
    lib.vecsz_decompress_host.argtypes = [vp, P, vp, u64, vp, u64]
    lib.vecsz_archive_info.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_int), u64p, ctypes.POINTER(ctypes.c_double),
                                       u64p, u64p]
    lib.vecsz_huffman_codebook.argtypes = [vp, ctypes.c_int, vp, vp, vp, u64, u64p, ctypes.POINTER(ctypes.c_int)]
    lib.vecsz_compress_batch.argtypes = [vp, ctypes.c_int, P, vp, vp, u64p, u64p, u64p, ctypes.POINTER(ctypes.c_int)]
    lib.vecsz_decompress_batch.argtypes = [vp, ctypes.c_int, P, vp, u64p, vp, u64p, ctypes.POINTER(ctypes.c_int)]
    lib.vecsz_ctx_set_workers.argtypes = [vp, ctypes.c_int]
    lib.vecsz_ctx_set_batch_fused.argtypes = [vp, ctypes.c_int]
    lib.vecsz_ctx_set_fused_encode.argtypes = [vp, ctypes.c_int]
    lib.vecsz_ctx_set_device_codebook.argtypes = [vp, ctypes.c_int]
    lib.vecsz_ctx_set_batch_graph.argtypes = [vp, ctypes.c_int]
    lib.vecsz_index_words.argtypes = [P, u64p]
    lib.vecsz_compress_indexed.argtypes = [vp, P, vp, vp, u64, u64p, u64p, vp, u64]
    lib.vecsz_decompress_indexed.argtypes = [vp, P, vp, u64, vp, u64, vp, u64]
    lib.vecsz_autotune.argtypes = [vp, P, vp, vp, u64, ctypes.POINTER(Tune)]
    lib.vecsz_ctx_set_launch_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.vecsz_server_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, u64]
    lib.vecsz_server_destroy.argtypes = [vp]
    lib.vecsz_server_submit_compress.argtypes = [vp, P, vp, vp, u64, u64p]
    lib.vecsz_server_submit_decompress.argtypes = [vp, P, vp, u64, vp, u64, u64p]
    lib.vecsz_server_wait.argtypes = [vp, u64, ctypes.POINTER(ServerResult)]
    lib.vecsz_lossless_bound.argtypes = [ctypes.c_int, u64, u64p]
    lib.vecsz_lossless_compress.argtypes = [ctypes.c_int, vp, u64, vp, u64, u64p]
    lib.vecsz_lossless_kind.argtypes = [vp, u64]
    lib.vecsz_lossless_decompress.argtypes = [vp, u64, vp, u64, u64p]
    lib.vecsz_sz_pack.argtypes = [vp, u64, u64, u64, vp, u64, u64p]
    lib.vecsz_sz_unpack.argtypes = [vp, u64, u64p, u64p, u64p, u64p]
    lib.vecsz_ctx_last_sidecar.argtypes = [vp, ctypes.POINTER(Sidecar)]
    lib.vecsz_params_from_sidecar.argtypes = [ctypes.POINTER(Sidecar), P]
    lib.vecsz_verify.argtypes = [vp, vp, vp, u64, ctypes.POINTER(Stat)]
    lib.vecsz_ctx_enable_timing.argtypes = [vp, ctypes.c_int]
    lib.vecsz_ctx_timings.argtypes = [vp, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double),
                                      ctypes.c_int]
    lib.vecsz_version.restype = ctypes.c_char_p
    _lib = lib
    return lib
