"""The one-pass encode (k_encode + k_enc_finish: bit counts, a decoupled
look-back over partitions, packing; opt-in) against the three-kernel encode
(chunk bits, scan, k_pack; the default) and the oracle: identical archives and decode indexes on
fields that exercise partition seams, ragged last partitions, outliers in
many partitions, a one-symbol tree, all-outlier fields, codewords longer than
32 bits (the packed LDS codebook with global escapes) and output buffers of
exactly the archive's size and one byte less."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _fib_deltas(nsym=35, seed=0):
    """1D field whose quant codes follow Fibonacci counts: the reference's
    heap then gives codewords of more than 32 bits.  Integer values at abs eb
    0.5 (prequant is the identity); each 8-block starts at 0 and steps by
    deltas 1..nsym, delta d occurring fib(nsym + 1 - d) times (the largest
    count on the smallest delta, which the heap's insertion order turns into
    the deepest tree)."""
    fib = [1, 1]
    while len(fib) < nsym:
        fib.append(fib[-1] + fib[-2])
    d = np.repeat(np.arange(1, nsym + 1), fib[::-1])
    rng = np.random.default_rng(seed)
    rng.shuffle(d)
    nblk = -(-d.size // 7)
    d = np.concatenate([d, np.ones(nblk * 7 - d.size, np.int64)]).reshape(nblk, 7)
    x = np.zeros((nblk, 8), np.float64)
    x[:, 1:] = np.cumsum(d, axis=1)
    return x.reshape(-1).astype(np.float32)


def _cases():
    from vecsz_amd import datagen as dg

    yield "grf48x64x80_r2r1e-4", dg.grf((48, 64, 80), seed=1), 1e-4, "r2r", {}
    yield "grf96x128x160_r2r1e-3", dg.grf((96, 128, 160), seed=2), 1e-3, "r2r", {}
    yield "sin300x517_abs1e-4", dg.sinusoids_2d(300, 517, seed=3), 1e-4, "abs", {}
    yield "walk1M+77_r2r1e-4", dg.walk_spikes_1d((1 << 20) + 77, seed=4), 1e-4, "r2r", {}
    yield "walk33k_b16", dg.walk_spikes_1d(33 * 1024 + 5, seed=5), 1e-4, "r2r", {"block": 16}
    yield "all_outliers", (np.random.default_rng(6).standard_normal((10, 11, 12)) * 100).astype(np.float32), \
        1e-6, "abs", {}
    yield "one_symbol", np.zeros((50, 60), np.float32), 1e-3, "abs", {}
    yield "grf40x61x75_dict1024", dg.grf((40, 61, 75), seed=7), 1e-3, "abs", {"dict_size": 1024}
    yield "fib_long_codes", _fib_deltas(), 0.5, "abs", {}


@pytest.mark.parametrize("name,x,eb,mode,kw", list(_cases()), ids=[c[0] for c in _cases()])
def test_one_pass_encode_matches_three_kernel_and_oracle(gpu_ctx, name, x, eb, mode, kw):
    print(f"encode {name} {x.shape}", flush=True)
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    a3, n3, i3 = gpu_ctx.compress_indexed(xd, eb, mode, **kw)
    try:
        gpu_ctx.set_fused_encode(True)
        a1, n1, i1 = gpu_ctx.compress_indexed(xd, eb, mode, **kw)
    finally:
        gpu_ctx.set_fused_encode(False)
    assert n1 == n3 and torch.equal(a1, a3), name
    assert torch.equal(i1, i3), name
    want, nout, _ = po.compress(x, eb, mode == "r2r", kw.get("dict_size", 4096), kw.get("block", 8))
    assert a1.cpu().numpy().tobytes() == want and n1 == nout
    if name == "fib_long_codes":
        from vecsz_amd import huffman_codebook

        codes, _ = po.dualquant(x, eb, False)[:2]
        _, lens, _, _ = huffman_codebook(np.bincount(codes.reshape(-1), minlength=4096).astype(np.uint64), 4096)
        assert int(lens.max()) > 32  # the packed LDS codebook with long-code escapes


def test_one_pass_encode_output_buffer_sizes(gpu_ctx):
    from vecsz_amd import VecszError, datagen as dg, make_params

    x = torch.from_numpy(dg.walk_spikes_1d(200_003, seed=8)).cuda()
    a, _ = gpu_ctx.compress(x, 1e-4, "r2r")
    p = make_params((x.numel(),), 1e-4, "r2r")
    try:
        gpu_ctx.set_fused_encode(True)
        for cap in (a.numel(), a.numel() - 1, a.numel() - 4 * 20, a.numel() // 2):
            out = torch.zeros(cap, dtype=torch.uint8, device="cuda")
            if cap == a.numel():
                size, _ = gpu_ctx.compress_into(p, x.data_ptr(), out.data_ptr(), cap)
                assert size == cap and torch.equal(out, a)
            else:
                with pytest.raises(VecszError):
                    gpu_ctx.compress_into(p, x.data_ptr(), out.data_ptr(), cap)
    finally:
        gpu_ctx.set_fused_encode(False)


@pytest.mark.parametrize("name,x,eb,mode,kw", list(_cases()), ids=[c[0] for c in _cases()])
def test_device_codebook_matches_host_codebook(gpu_ctx, name, x, eb, mode, kw):
    """The codebook built on the device (the reference's queue replayed by one
    lane, codes and tree bytes in parallel; the compress then has no stop on
    the host) gives the host build's archive, decode index and sidecar."""
    print(f"device codebook {name} {x.shape}", flush=True)
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    a0, n0, i0 = gpu_ctx.compress_indexed(xd, eb, mode, **kw)
    s0 = gpu_ctx.last_sidecar()
    try:
        gpu_ctx.set_device_codebook(True)
        a1, n1, i1 = gpu_ctx.compress_indexed(xd, eb, mode, **kw)
        s1 = gpu_ctx.last_sidecar()
    finally:
        gpu_ctx.set_device_codebook(False)
    assert n1 == n0 and torch.equal(a1, a0), name
    assert torch.equal(i1, i0) and s1 == s0, name


def test_device_codebook_errors(gpu_ctx):
    from vecsz_amd import VecszError, datagen as dg, make_params

    x = torch.from_numpy(dg.grf((20, 30, 40), seed=9)).cuda()
    a, _ = gpu_ctx.compress(x, 1e-4, "r2r")
    p = make_params((40, 30, 20), 1e-4, "r2r")
    try:
        gpu_ctx.set_device_codebook(True)
        for cap in (a.numel() - 1, 200, 40):  # short of the stream, of the tree, of the header
            out = torch.zeros(cap, dtype=torch.uint8, device="cuda")
            with pytest.raises(VecszError):
                gpu_ctx.compress_into(p, x.data_ptr(), out.data_ptr(), cap)
        with pytest.raises(VecszError):  # zero r2r range, found at the end of the call
            gpu_ctx.compress(torch.zeros(64, 64, device="cuda"), 1e-3, "r2r")
        out = torch.zeros(a.numel(), dtype=torch.uint8, device="cuda")
        size, _ = gpu_ctx.compress_into(p, x.data_ptr(), out.data_ptr(), a.numel())
        assert size == a.numel() and torch.equal(out, a)
    finally:
        gpu_ctx.set_device_codebook(False)
