"""MI355X (gfx950) implementation of vecSZ's dual-quantization + Huffman
compression path behind the reference's Compress/Decompress interface.

The work runs in libvecsz_hip.so (hand-written HIP kernels + C ABI, see
include/vecsz_hip.h); this package is the host-side mirror and its loader.
"""
from .interface import ArgParse, Compress, Context, Decompress, Dim4, Server, archive_info, huffman_codebook, lossless_compress, lossless_decompress, lossless_kind, make_params, params_from_sidecar, sz_pack, sz_unpack  # noqa: F401
from ._lib import VecszError, LIB_PATH  # noqa: F401

__all__ = ["ArgParse", "Compress", "Decompress", "Context", "Server", "Dim4", "archive_info", "huffman_codebook", "make_params", "VecszError"]
