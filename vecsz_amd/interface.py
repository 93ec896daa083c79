"""Host-side mirror of the reference's `vecsz::interface` (src/vecsz_interface.hh).

`Compress(ap, data_in)` and `Decompress(ap, archive)` keep the reference's
names and argument meaning: `ap` carries the fields of the reference's
ArgParse that the path reads (ndim, dim4, eb, mode, dict_size, block_size),
and the results are the archive payload (datapack::pack layout) and the
reconstructed field.  They run on the GPU through libvecsz_hip.so; there is no
CPU fallback, and a missing library raises.

`Context` is the lower-level API: device tensors in, device tensors out, on
one HIP stream.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import Params, VecszError

try:  # torch supplies device buffers and the stream; the library does the work
    import torch
except Exception:  # pragma: no cover
    torch = None


def _mode_id(mode) -> int:
    if isinstance(mode, int):
        return mode
    m = str(mode).lower()
    if m == "abs":
        return _lib.MODE_ABS
    if m in ("r2r", "vrrel"):
        return _lib.MODE_R2R
    raise ValueError(f"mode must be 'abs' or 'r2r', got {mode!r}")


def make_params(dims_xfirst: Sequence[int], eb: float, mode="abs", block: int = 8, dict_size: int = 4096,
                value_range: float = 0.0, pad_mode="none", pad_type="const", pad_constant: float = 0.0) -> Params:
    """Compress / decompress parameters.  pad_mode 'none' | 'global' | 'block'
    | 'edge' and pad_type 'max' | 'min' | 'avg' | 'const' select the
    reference's padding (--padding global,max etc.); decompress reads the
    padding from the archive."""
    p = Params()
    d = [int(v) for v in dims_xfirst]
    if not 1 <= len(d) <= 3:
        raise ValueError("1 to 3 dimensions")
    p.ndim = len(d)
    for i in range(3):
        p.dims[i] = d[i] if i < len(d) else 1
    p.mode = _mode_id(mode)
    p.eb = float(eb)
    p.value_range = float(value_range)
    p.block = int(block)
    p.dict_size = int(dict_size)
    p.pad_mode = _lib.PAD_MODES[pad_mode] if isinstance(pad_mode, str) else int(pad_mode)
    p.pad_type = _lib.PAD_TYPES[pad_type] if isinstance(pad_type, str) else int(pad_type)
    p.pad_constant = float(pad_constant)
    return p


def dims_of(shape) -> tuple:
    """C-order shape (slowest first) -> archive dims (x first)."""
    return tuple(int(s) for s in reversed(tuple(shape)))


def archive_info(archive) -> dict:
    lib = _lib.load()
    b = bytes(archive) if not isinstance(archive, (bytes, bytearray)) else bytes(archive)
    nd = ctypes.c_int()
    dims = (ctypes.c_uint64 * 3)()
    eb = ctypes.c_double()
    hs = ctypes.c_uint64()
    no = ctypes.c_uint64()
    rc = lib.vecsz_archive_info(b, len(b), ctypes.byref(nd), dims, ctypes.byref(eb), ctypes.byref(hs), ctypes.byref(no))
    if rc:
        raise VecszError(rc, lib.vecsz_strerror(rc).decode())
    return {"ndim": nd.value, "dims": tuple(dims[i] for i in range(nd.value)), "eb": eb.value,
            "huff_size": hs.value, "n_outlier": no.value}


def huffman_codebook(freq, dict_size: Optional[int] = None):
    """The reference's Huffman codebook for a quant-code histogram
    (src/huffman.hh:239-262, :200-227, :425-502), built on the host by the
    library: (code_bits u64[dict], code_len u8[dict], node_count, tree bytes)."""
    lib = _lib.load()
    f = np.ascontiguousarray(freq, dtype=np.uint64)
    d = int(dict_size if dict_size is not None else f.size)
    if f.size < d:
        raise ValueError("histogram shorter than dict_size")
    bits = np.zeros(d, dtype=np.uint64)
    lens = np.zeros(d, dtype=np.uint8)
    tb = ctypes.c_uint64()
    nc = ctypes.c_int()
    args = (f.ctypes.data_as(ctypes.c_void_p), d, bits.ctypes.data_as(ctypes.c_void_p),
            lens.ctypes.data_as(ctypes.c_void_p))
    rc = lib.vecsz_huffman_codebook(*args, None, 0, ctypes.byref(tb), ctypes.byref(nc))
    if rc:
        raise VecszError(rc, lib.vecsz_strerror(rc).decode())
    tree = np.zeros(tb.value, dtype=np.uint8)
    rc = lib.vecsz_huffman_codebook(*args, tree.ctypes.data_as(ctypes.c_void_p), tb.value, ctypes.byref(tb),
                                    ctypes.byref(nc))
    if rc:
        raise VecszError(rc, lib.vecsz_strerror(rc).decode())
    return bits, lens, nc.value, tree.tobytes()


LOSSLESS = {"gzip": _lib.VECSZ_LOSSLESS_GZIP if hasattr(_lib, "VECSZ_LOSSLESS_GZIP") else 1,
            "zstd": _lib.VECSZ_LOSSLESS_ZSTD if hasattr(_lib, "VECSZ_LOSSLESS_ZSTD") else 2}


def _buf(b) -> np.ndarray:
    return np.frombuffer(bytes(b), dtype=np.uint8)


def lossless_compress(kind: str, payload) -> bytes:
    """The reference's lossless pass (gzip or zstd, level 3) over a payload
    (src/lossless.hh:179-198)."""
    lib = _lib.load()
    k = LOSSLESS[kind]
    src = _buf(payload)
    bound = ctypes.c_uint64()
    rc = lib.vecsz_lossless_bound(k, src.size, ctypes.byref(bound))
    if rc:
        raise VecszError(rc, f"{kind} pass unavailable")
    out = np.empty(bound.value, dtype=np.uint8)
    n = ctypes.c_uint64()
    rc = lib.vecsz_lossless_compress(k, src.ctypes.data_as(ctypes.c_void_p), src.size,
                                     out.ctypes.data_as(ctypes.c_void_p), out.size, ctypes.byref(n))
    if rc:
        raise VecszError(rc, lib.vecsz_strerror(rc).decode())
    return out[: n.value].tobytes()


def lossless_kind(data) -> Optional[str]:
    src = _buf(data[:64])
    k = _lib.load().vecsz_lossless_kind(src.ctypes.data_as(ctypes.c_void_p), src.size)
    return {1: "gzip", 2: "zstd"}.get(k)


def lossless_decompress(data, target: int) -> bytes:
    """Inverse of the lossless pass: the `target`-byte payload
    (src/lossless.hh:200-218)."""
    lib = _lib.load()
    src = _buf(data)
    out = np.empty(int(target), dtype=np.uint8)
    n = ctypes.c_uint64()
    rc = lib.vecsz_lossless_decompress(src.ctypes.data_as(ctypes.c_void_p), src.size,
                                       out.ctypes.data_as(ctypes.c_void_p), int(target), ctypes.byref(n))
    if rc:
        raise VecszError(rc, lib.vecsz_strerror(rc).decode())
    return out.tobytes()


def sz_pack(body, lossless_size: int, data_size: int) -> bytes:
    """.sz file bytes (src/vecsz.cc:43-56): header + body."""
    lib = _lib.load()
    src = _buf(body)
    out = np.empty(src.size + 16, dtype=np.uint8)
    n = ctypes.c_uint64()
    rc = lib.vecsz_sz_pack(src.ctypes.data_as(ctypes.c_void_p), src.size, int(lossless_size), int(data_size),
                           out.ctypes.data_as(ctypes.c_void_p), out.size, ctypes.byref(n))
    if rc:
        raise VecszError(rc, lib.vecsz_strerror(rc).decode())
    return out[: n.value].tobytes()


def sz_unpack(file_bytes) -> tuple:
    """(body, lossless_size, data_size) of .sz file bytes."""
    lib = _lib.load()
    src = _buf(file_bytes)
    ls, ds, off, nb = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib.vecsz_sz_unpack(src.ctypes.data_as(ctypes.c_void_p), src.size, ctypes.byref(ls), ctypes.byref(ds),
                             ctypes.byref(off), ctypes.byref(nb))
    if rc:
        raise VecszError(rc, lib.vecsz_strerror(rc).decode())
    return src[off.value: off.value + nb.value].tobytes(), ls.value, ds.value


def params_from_sidecar(record: bytes) -> Params:
    """Decompress parameters from a metadata record (Context.last_sidecar)."""
    sc = _lib.Sidecar.from_buffer_copy(bytes(record))
    p = Params()
    rc = _lib.load().vecsz_params_from_sidecar(ctypes.byref(sc), ctypes.byref(p))
    if rc:
        raise VecszError(rc, "not a vecsz metadata record")
    return p


def header_fields(head: bytes) -> dict:
    """ndim, dims (x first) and eb from the first bytes of an archive payload
    (datapack layout, src/utils/datapack.hh:33-47); no section-size checks."""
    nd = int(np.frombuffer(head[:4], dtype=np.int32)[0])
    if not 1 <= nd <= 3 or len(head) < 4 * (nd + 1) + 8:
        raise VecszError(_lib.VECSZ_ECORRUPT, "archive header is corrupt")
    dims = tuple(int(v) for v in np.frombuffer(head[4:4 + 4 * nd], dtype=np.int32))
    if min(dims) <= 0:
        raise VecszError(_lib.VECSZ_ECORRUPT, "archive header is corrupt")
    eb = float(np.frombuffer(head[4 * (nd + 1):4 * (nd + 1) + 8], dtype=np.float64)[0])
    return {"ndim": nd, "dims": dims, "eb": eb}


class Context:
    """One GPU, one HIP stream, reusable device scratch."""

    def __init__(self, device: int = 0, stream=None):
        self._lib = _lib.load()
        self.device = device
        handle = None
        if stream is None and torch is not None and torch.cuda.is_available():
            stream = torch.cuda.current_stream(device)
        if stream is not None:
            handle = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)
        self.stream = stream
        ctx = ctypes.c_void_p()
        rc = self._lib.vecsz_ctx_create(ctypes.byref(ctx), device, ctypes.c_void_p(handle) if handle else None)
        if rc:
            raise VecszError(rc, self._lib.vecsz_strerror(rc).decode())
        self._ctx = ctx

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.vecsz_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            msg = self._lib.vecsz_ctx_last_error(self._ctx).decode()
            raise VecszError(rc, msg or self._lib.vecsz_strerror(rc).decode())

    def enable_timing(self, on: bool = True):
        self._check(self._lib.vecsz_ctx_enable_timing(self._ctx, int(on)))

    def timings(self) -> list:
        names = (ctypes.c_char_p * 64)()
        ms = (ctypes.c_double * 64)()
        n = self._lib.vecsz_ctx_timings(self._ctx, names, ms, 64)
        return [(names[i].decode(), ms[i]) for i in range(n)]

    # -------------------------------------------------------------- device API

    def last_sidecar(self) -> bytes:
        """The 48-byte metadata record (mode, B, dict, r2r range, eb_final) of
        this context's last compress: keep it beside the archive and
        decompress needs nothing else."""
        sc = _lib.Sidecar()
        self._check(self._lib.vecsz_ctx_last_sidecar(self._ctx, ctypes.byref(sc)))
        return bytes(sc)

    def verify(self, orig, recon) -> dict:
        """VerifyData (src/utils/verify.hh:30-74) on the device for two float32
        CUDA tensors of the same size."""
        if orig.numel() != recon.numel():
            raise ValueError("size mismatch")
        o, x = orig.contiguous(), recon.contiguous()
        st = _lib.Stat()
        self._check(self._lib.vecsz_verify(self._ctx, ctypes.c_void_p(o.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                           o.numel(), ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def set_workers(self, n: int):
        """Worker contexts of the batched calls (1..16)."""
        self._check(self._lib.vecsz_ctx_set_workers(self._ctx, int(n)))

    def autotune(self, x, eb: float, mode="abs", block: int = 8, dict_size: int = 4096, value_range: float = 0.0):
        """Launch-shape autotuner (src/autotune.hh:80-127 extended to GPU
        launch shapes) on a float32 CUDA tensor: keeps the fastest dual-quant
        and encode workgroup caps on this context and reports, per Lorenzo
        block size, the whole-compress time and archive size (B itself is
        not applied)."""
        if not x.is_cuda or x.dtype != torch.float32:
            raise TypeError("expected a float32 CUDA tensor")
        x = x.contiguous()
        p = make_params(dims_of(x.shape), eb, mode, block, dict_size, value_range)
        q = Params()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(p))
        q.block = 64
        cap = max(self.compress_bound(p), self.compress_bound(q))
        out = torch.empty(cap, dtype=torch.uint8, device=x.device)
        t = _lib.Tune()
        self._check(self._lib.vecsz_autotune(self._ctx, ctypes.byref(p), ctypes.c_void_p(x.data_ptr()),
                                             ctypes.c_void_p(out.data_ptr()), cap, ctypes.byref(t)))
        return {"block": t.block, "dq_grid": t.dq_grid, "enc_grid": t.enc_grid, "dq_ms": t.dq_ms,
                "enc_ms": t.enc_ms, "ms_by_block": dict(zip((8, 16, 32, 64), list(t.ms_by_block))),
                "bytes_by_block": dict(zip((8, 16, 32, 64), [int(b) for b in t.bytes_by_block]))}

    def set_launch_shape(self, dq_grid: int, enc_grid: int):
        """Workgroup caps of the dual-quantization and encode kernels (what
        autotune() picks); archives do not depend on them."""
        self._check(self._lib.vecsz_ctx_set_launch_shape(self._ctx, int(dq_grid), int(enc_grid)))

    def set_device_codebook(self, on: bool):
        """Codebook built on the device (True: no host stop) or on the host
        (False, the default: faster); identical archives."""
        self._check(self._lib.vecsz_ctx_set_device_codebook(self._ctx, 1 if on else 0))

    def set_fused_encode(self, on: bool):
        """One-pass encode (True) or the three-kernel encode (False, the
        default: faster on an MI355X); identical archives."""
        self._check(self._lib.vecsz_ctx_set_fused_encode(self._ctx, 1 if on else 0))

    def set_batch_fused(self, on):
        """True (default): fused batches (one launch per stage for all
        fields) when they qualify, decompress batches only when their fields
        are small on average; "always": decompress batches of large fields
        fused as well; False: the worker-context path."""
        mode = 2 if on == "always" else 1 if on else 0
        self._check(self._lib.vecsz_ctx_set_batch_fused(self._ctx, mode))

    def set_batch_graph(self, on: bool):
        """True (default): a fused compress batch whose layout repeats runs
        as two captured HIP graphs (before and after its host codebook
        stop); False: always eagerly.  Identical archives."""
        self._check(self._lib.vecsz_ctx_set_batch_graph(self._ctx, 1 if on else 0))

    def compress_bound(self, params: Params) -> int:
        b = ctypes.c_uint64()
        self._check(self._lib.vecsz_compress_bound(ctypes.byref(params), ctypes.byref(b)))
        return b.value

    def compress_into(self, params: Params, d_in_ptr: int, d_out_ptr: int, out_cap: int, index_ptr: int = 0,
                      index_words: int = 0):
        """Raw device pointers; returns (archive bytes, n_outlier).  With
        index_ptr, also writes the decode index (index_words u64 words)."""
        size = ctypes.c_uint64()
        nout = ctypes.c_uint64()
        if index_ptr:
            rc = self._lib.vecsz_compress_indexed(self._ctx, ctypes.byref(params), ctypes.c_void_p(d_in_ptr),
                                                  ctypes.c_void_p(d_out_ptr), out_cap, ctypes.byref(size),
                                                  ctypes.byref(nout), ctypes.c_void_p(index_ptr), index_words)
        else:
            rc = self._lib.vecsz_compress(self._ctx, ctypes.byref(params), ctypes.c_void_p(d_in_ptr),
                                          ctypes.c_void_p(d_out_ptr), out_cap, ctypes.byref(size), ctypes.byref(nout))
        self._check(rc)
        return size.value, nout.value

    def decompress_into(self, params: Params, d_in_ptr: int, in_bytes: int, d_out_ptr: int, out_cap: int,
                        index_ptr: int = 0, index_words: int = 0):
        if index_ptr:
            rc = self._lib.vecsz_decompress_indexed(self._ctx, ctypes.byref(params), ctypes.c_void_p(d_in_ptr),
                                                    in_bytes, ctypes.c_void_p(index_ptr), index_words,
                                                    ctypes.c_void_p(d_out_ptr), out_cap)
        else:
            rc = self._lib.vecsz_decompress(self._ctx, ctypes.byref(params), ctypes.c_void_p(d_in_ptr), in_bytes,
                                            ctypes.c_void_p(d_out_ptr), out_cap)
        self._check(rc)

    def index_words(self, params: Params) -> int:
        """u64 words of the decode index for a field of these dims."""
        w = ctypes.c_uint64()
        self._check(self._lib.vecsz_index_words(ctypes.byref(params), ctypes.byref(w)))
        return w.value

    def compress_indexed(self, x, eb: float = 1e-4, mode="abs", block: int = 8, dict_size: int = 4096,
                         value_range: float = 0.0, pad_mode="none", pad_type="const", pad_constant: float = 0.0):
        """compress() of a float32 CUDA tensor that also returns the decode
        index: (archive, n_outlier, index as an int64 CUDA tensor)."""
        if not x.is_cuda or x.dtype != torch.float32:
            raise TypeError("expected a float32 CUDA tensor")
        x = x.contiguous()
        p = make_params(dims_of(x.shape), eb, mode, block, dict_size, value_range, pad_mode, pad_type, pad_constant)
        cap = self.compress_bound(p)
        out = torch.empty(cap, dtype=torch.uint8, device=x.device)
        idx = torch.empty(self.index_words(p), dtype=torch.int64, device=x.device)
        size, nout = self.compress_into(p, x.data_ptr(), out.data_ptr(), out.numel(), idx.data_ptr(), idx.numel())
        return out[:size], nout, idx

    def compress_batch_into(self, params, d_in, d_out, caps):
        """Raw device pointers, one entry per field; returns [(bytes, n_outlier)]."""
        nf = len(params)
        P = (Params * nf)(*params)
        vp = ctypes.c_void_p
        ins = (vp * nf)(*[vp(int(x)) for x in d_in])
        outs = (vp * nf)(*[vp(int(x)) for x in d_out])
        cap = (ctypes.c_uint64 * nf)(*[int(c) for c in caps])
        sizes = (ctypes.c_uint64 * nf)()
        nout = (ctypes.c_uint64 * nf)()
        st = (ctypes.c_int * nf)()
        self._check(self._lib.vecsz_compress_batch(self._ctx, nf, P, ins, outs, cap, sizes, nout, st))
        return [(sizes[k], nout[k]) for k in range(nf)]

    def decompress_batch_into(self, params, d_in, in_bytes, d_out, caps, check: bool = True):
        """Raw device pointers, one entry per field.  Raises on failure; with
        check=False returns (status of the call, [status of each field])."""
        nf = len(params)
        P = (Params * nf)(*params)
        vp = ctypes.c_void_p
        ins = (vp * nf)(*[vp(int(x)) for x in d_in])
        nb = (ctypes.c_uint64 * nf)(*[int(b) for b in in_bytes])
        outs = (vp * nf)(*[vp(int(x)) for x in d_out])
        cap = (ctypes.c_uint64 * nf)(*[int(c) for c in caps])
        st = (ctypes.c_int * nf)()
        rc = self._lib.vecsz_decompress_batch(self._ctx, nf, P, ins, nb, outs, cap, st)
        if check:
            self._check(rc)
        return rc, [st[k] for k in range(nf)]

    def compress_batch(self, fields, ebs, modes="abs", block: int = 8, dict_size: int = 4096):
        """Compress a list of float32 CUDA tensors in one batched call;
        returns [(uint8 CUDA tensor, n_outlier)], each identical to compress()."""
        nf = len(fields)
        ebs = list(ebs) if isinstance(ebs, (list, tuple)) else [ebs] * nf
        modes = list(modes) if isinstance(modes, (list, tuple)) else [modes] * nf
        fields = [f.contiguous() for f in fields]
        params, outs = [], []
        for f, eb, m in zip(fields, ebs, modes):
            if not f.is_cuda or f.dtype != torch.float32:
                raise TypeError("expected float32 CUDA tensors")
            p = make_params(dims_of(f.shape), eb, m, block, dict_size)
            params.append(p)
            outs.append(torch.empty(self.compress_bound(p), dtype=torch.uint8, device=f.device))
        res = self.compress_batch_into(params, [f.data_ptr() for f in fields], [o.data_ptr() for o in outs],
                                       [o.numel() for o in outs])
        return [(o[:size], nout) for o, (size, nout) in zip(outs, res)]

    def decompress_batch(self, archives, modes="abs", value_ranges=0.0, block: int = 8, dict_size: int = 4096):
        """Decompress a list of uint8 CUDA archives in one batched call."""
        nf = len(archives)
        modes = list(modes) if isinstance(modes, (list, tuple)) else [modes] * nf
        vrs = list(value_ranges) if isinstance(value_ranges, (list, tuple)) else [value_ranges] * nf
        params, outs = [], []
        for a, m, vr in zip(archives, modes, vrs):
            info = header_fields(a[:64].cpu().numpy().tobytes())
            shape = tuple(reversed(info["dims"]))
            params.append(make_params((1,), 1e-4, m, block, dict_size, vr))
            outs.append(torch.empty(shape, dtype=torch.float32, device=a.device))
        self.decompress_batch_into(params, [a.data_ptr() for a in archives], [a.numel() for a in archives],
                                   [o.data_ptr() for o in outs], [o.numel() for o in outs])
        return outs

    def compress(self, x, eb: float = 1e-4, mode="abs", block: int = 8, dict_size: int = 4096,
                 value_range: float = 0.0, out=None, pad_mode="none", pad_type="const", pad_constant: float = 0.0):
        """Compress a float32 field.

        torch CUDA tensor -> (uint8 CUDA tensor holding the archive, n_outlier);
        numpy array       -> (bytes, n_outlier) through host staging."""
        pad = dict(pad_mode=pad_mode, pad_type=pad_type, pad_constant=pad_constant)
        if torch is not None and isinstance(x, torch.Tensor):
            if not x.is_cuda or x.dtype != torch.float32:
                raise TypeError("expected a float32 CUDA tensor")
            x = x.contiguous()
            p = make_params(dims_of(x.shape), eb, mode, block, dict_size, value_range, **pad)
            cap = self.compress_bound(p)
            if out is None or out.numel() < cap:
                out = torch.empty(cap, dtype=torch.uint8, device=x.device)
            size, nout = self.compress_into(p, x.data_ptr(), out.data_ptr(), out.numel())
            return out[:size], nout
        a = np.ascontiguousarray(x, dtype=np.float32)
        p = make_params(dims_of(a.shape), eb, mode, block, dict_size, value_range, **pad)
        cap = self.compress_bound(p)
        buf = np.empty(cap, dtype=np.uint8)
        size = ctypes.c_uint64()
        nout = ctypes.c_uint64()
        rc = self._lib.vecsz_compress_host(self._ctx, ctypes.byref(p), a.ctypes.data_as(ctypes.c_void_p),
                                           buf.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(size),
                                           ctypes.byref(nout))
        self._check(rc)
        return buf[:size.value].tobytes(), nout.value

    def decompress(self, archive, mode="abs", value_range: float = 0.0, block: int = 8, dict_size: int = 4096,
                   out=None, index=None):
        """Decompress an archive payload.

        uint8 CUDA tensor -> float32 CUDA tensor (C-order shape from the archive);
        bytes / numpy     -> numpy array.  index: the decode index that
        compress_indexed() returned with this archive (CUDA archives only)."""
        p = make_params((1,), 1e-4, mode, block, dict_size, value_range)
        if torch is not None and isinstance(archive, torch.Tensor):
            if not archive.is_cuda or archive.dtype != torch.uint8:
                raise TypeError("expected a uint8 CUDA tensor")
            head = archive[:64].cpu().numpy().tobytes()
            info = header_fields(head)
            shape = tuple(reversed(info["dims"]))
            n = int(np.prod(shape))
            if out is None:
                out = torch.empty(shape, dtype=torch.float32, device=archive.device)
            if index is not None:
                self.decompress_into(p, archive.data_ptr(), archive.numel(), out.data_ptr(), out.numel(),
                                     index.data_ptr(), index.numel())
            else:
                self.decompress_into(p, archive.data_ptr(), archive.numel(), out.data_ptr(), out.numel())
            return out.view(shape) if out.numel() == n else out
        b = bytes(archive)
        info = archive_info(b)
        shape = tuple(reversed(info["dims"]))
        n = int(np.prod(shape))
        y = np.empty(n, dtype=np.float32)
        rc = self._lib.vecsz_decompress_host(self._ctx, ctypes.byref(p), b, len(b), y.ctypes.data_as(ctypes.c_void_p), n)
        self._check(rc)
        return y.reshape(shape)


class Server:
    """Request server on one GPU (vecsz_server_*): submit compress and
    decompress requests on device buffers, wait for each by ticket.  Requests
    overlap on the GPU across `slots` slot contexts; admission keeps the bytes
    held by running requests under `hbm_budget`."""

    def __init__(self, device: int = 0, slots: int = 8, hbm_budget: int = 200 * 10**9):
        self._lib = _lib.load()
        s = ctypes.c_void_p()
        rc = self._lib.vecsz_server_create(ctypes.byref(s), device, slots, int(hbm_budget))
        if rc:
            raise VecszError(rc, self._lib.vecsz_strerror(rc).decode())
        self._s = s

    def close(self):
        if getattr(self, "_s", None):
            self._lib.vecsz_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ticket(self, rc, t):
        if rc:
            raise VecszError(rc, self._lib.vecsz_strerror(rc).decode())
        return t.value

    def submit_compress(self, params: Params, d_in: int, d_out: int, out_cap: int) -> int:
        t = ctypes.c_uint64()
        rc = self._lib.vecsz_server_submit_compress(self._s, ctypes.byref(params), ctypes.c_void_p(d_in),
                                                    ctypes.c_void_p(d_out), out_cap, ctypes.byref(t))
        return self._ticket(rc, t)

    def submit_decompress(self, params: Params, d_in: int, in_bytes: int, d_out: int, out_cap: int) -> int:
        t = ctypes.c_uint64()
        rc = self._lib.vecsz_server_submit_decompress(self._s, ctypes.byref(params), ctypes.c_void_p(d_in), in_bytes,
                                                      ctypes.c_void_p(d_out), out_cap, ctypes.byref(t))
        return self._ticket(rc, t)

    def wait(self, ticket: int) -> dict:
        r = _lib.ServerResult()
        rc = self._lib.vecsz_server_wait(self._s, ticket, ctypes.byref(r))
        if rc:
            raise VecszError(rc, self._lib.vecsz_strerror(rc).decode())
        if r.status:
            raise VecszError(r.status, self._lib.vecsz_strerror(r.status).decode())
        return {"out_bytes": r.out_bytes, "n_outlier": r.n_outlier, "queue_ms": r.queue_ms, "run_ms": r.run_ms,
                "latency_ms": r.latency_ms}


# ---------------------------------------------------------------- reference-shaped API

@dataclass
class Dim4:
    _0: int = 1
    _1: int = 1
    _2: int = 1
    _3: int = 1


@dataclass
class ArgParse:
    """The ArgParse fields the path reads (src/argument_parser/argparse.hh:45-98),
    with the reference's defaults (:62-77).  `value_range` stands in for the
    reference's re-read of the input file in r2r mode."""
    ndim: int = -1
    dim4: Dim4 = field(default_factory=Dim4)
    eb: float = 1e-4
    mode: str = "abs"
    dict_size: int = 4096
    radius: int = 2048
    block_size: int = 8
    value_range: float = 0.0
    device: int = 0
    lossless: Optional[str] = None  # szwf.lossless_pass: None, "gzip" or "zstd" (argparse.hh:32-35)
    # szwf.{global,block,edge}_padding, pad_type (0 max, 1 min, 2 avg, 3 constant)
    # and pad_constant (argparse.hh:34-36, :75-77)
    global_padding: bool = False
    block_padding: bool = False
    edge_padding: bool = False
    pad_type: int = 3
    pad_constant: float = 0.0

    def pad_mode(self) -> str:
        """One padding mode at a time: global wins over block and edge, block
        over edge (src/argument_parser/argparse.cc:156-169)."""
        if self.global_padding:
            return "global"
        if self.block_padding:
            return "block"
        if self.edge_padding:
            return "edge"
        return "none"


_default_ctx: Optional[Context] = None


def _ctx(ap: ArgParse) -> Context:
    global _default_ctx
    if _default_ctx is None or _default_ctx.device != ap.device:
        _default_ctx = Context(ap.device)
    return _default_ctx


def Compress(ap: ArgParse, data_in):
    """vecsz::interface::Compress<float,int> (src/vecsz_interface.hh:31-229).
    Returns (archive, num_outlier, data_size, lossless_size): with
    ap.lossless set, the archive is the lossless stream of the data_size-byte
    payload (:206-217), else lossless_size is 0."""
    dims = [ap.dim4._0, ap.dim4._1, ap.dim4._2][: ap.ndim]
    if torch is not None and isinstance(data_in, torch.Tensor):
        x = data_in.reshape(tuple(reversed(dims)))
    else:
        x = np.asarray(data_in, dtype=np.float32).reshape(tuple(reversed(dims)))
    arch, nout = _ctx(ap).compress(x, ap.eb, ap.mode, ap.block_size, ap.dict_size, ap.value_range,
                                   pad_mode=ap.pad_mode(), pad_type=int(ap.pad_type), pad_constant=ap.pad_constant)
    size = arch.numel() if hasattr(arch, "numel") else len(arch)
    if ap.lossless:
        payload = arch.cpu().numpy().tobytes() if hasattr(arch, "numel") else arch
        packed = lossless_compress(ap.lossless, payload)
        return packed, nout, size, len(packed)
    return arch, nout, size, 0


def Decompress(ap: ArgParse, tmp_data, target_size: int = 0, lossless_size: int = 0):
    """vecsz::interface::Decompress<float,int> (src/vecsz_interface.hh:232-375).
    Sets ap.ndim / ap.dim4 / ap.eb from the archive like the reference does
    (:262-268) and returns the reconstructed field."""
    if lossless_size:
        # lossless pass first (:237-244): the stream inflates to the target_size payload
        raw = tmp_data.cpu().numpy().tobytes() if hasattr(tmp_data, "numel") else bytes(tmp_data)
        tmp_data = lossless_decompress(raw[:lossless_size], target_size)
    y = _ctx(ap).decompress(tmp_data, ap.mode, ap.value_range, ap.block_size, ap.dict_size)
    dims = tuple(reversed(tuple(y.shape)))
    ap.ndim = len(dims)
    ap.dim4 = Dim4(*(list(dims) + [1] * (4 - len(dims))))
    if torch is not None and isinstance(tmp_data, torch.Tensor):
        head = tmp_data[:64].cpu().numpy().tobytes()
    else:
        head = bytes(tmp_data[:64])
    ap.eb = header_fields(head)["eb"]
    return y
