"""The HIP path (through the C ABI) against the reference's golden outputs and
the C oracle: archives byte-identical, decompressed floats bit-identical."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _mode(c):
    return "r2r" if c["r2r"] else "abs"


def _say(*a):
    print(*a, flush=True)


def test_golden_compress_bit_exact(gpu_ctx, golden):
    bad = []
    for c in golden["cases"]:
        _say("compress", c["name"])
        x = torch.from_numpy(c["x"]).cuda()
        a, nout = gpu_ctx.compress(x, c["eb"], _mode(c), c["block"], c["dict"])
        if a.cpu().numpy().tobytes() != c["archive"] or nout != c["n_outlier"]:
            bad.append(c["name"])
    assert not bad, bad


def test_golden_decompress_bit_exact(gpu_ctx, golden):
    bad = []
    for c in golden["cases"]:
        _say("decompress", c["name"])
        a = torch.from_numpy(np.frombuffer(c["archive"], dtype=np.uint8).copy()).cuda()
        y = gpu_ctx.decompress(a, _mode(c), c["value_range"], c["block"], c["dict"])
        y = y.cpu().numpy().reshape(c["x"].shape)
        if not np.array_equal(y.view(np.uint32), c["out"].view(np.uint32)):
            bad.append(c["name"])
    assert not bad, bad


def test_host_entry_points(gpu_ctx, golden):
    for c in golden["cases"][:8]:
        _say("host", c["name"])
        a, nout = gpu_ctx.compress(c["x"], c["eb"], _mode(c), c["block"], c["dict"])
        assert a == c["archive"]
        y = gpu_ctx.decompress(a, _mode(c), c["value_range"], c["block"], c["dict"])
        assert np.array_equal(y.view(np.uint32), c["out"].view(np.uint32))


def test_reference_shaped_interface(golden):
    from vecsz_amd import ArgParse, Compress, Decompress, Dim4

    c = [c for c in golden["cases"] if c["name"] == "3d_grf_partial_x_1e-4"][0]
    ap = ArgParse(ndim=3, dim4=Dim4(*reversed(c["x"].shape)), eb=c["eb"], mode="r2r",
                  value_range=c["value_range"])
    arch, nout, size, lossless = Compress(ap, c["x"].ravel())
    assert arch == c["archive"] and size == len(c["archive"]) and lossless == 0
    ap2 = ArgParse(mode="r2r", value_range=c["value_range"])
    y = Decompress(ap2, arch, size, 0)
    assert ap2.ndim == 3 and ap2.eb == c["eb"]
    assert np.array_equal(y.view(np.uint32), c["out"].view(np.uint32))


def _check_vs_oracle(ctx, x, eb, r2r, block=8, dict_size=4096, tag="", offset=0):
    """offset: the device field starts `offset` floats into its allocation
    (offset 1: not 16-byte aligned)."""
    _say("oracle-compare", tag, x.shape, eb, r2r, block, offset)
    a_orc, n_orc, _ = po.compress(x, eb, r2r, dict_size, block)
    xt = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    if offset:
        buf = torch.zeros(x.size + offset, dtype=torch.float32, device="cuda")
        buf[offset:] = xt.ravel()
        xt = buf[offset:].view(x.shape)
    a, nout = ctx.compress(xt, eb, "r2r" if r2r else "abs", block, dict_size)
    ab = a.cpu().numpy().tobytes()
    assert nout == n_orc, tag
    assert len(ab) == len(a_orc), tag
    assert ab == a_orc, tag
    rng = po.value_range(x) if r2r else 0.0
    y_orc = po.decompress(a_orc, x.shape, r2r, rng, dict_size, block)
    y = ctx.decompress(a, "r2r" if r2r else "abs", rng, block, dict_size).cpu().numpy().reshape(x.shape)
    assert np.array_equal(y.view(np.uint32), y_orc.view(np.uint32)), tag


@pytest.mark.parametrize("shape", [(1,), (7,), (4097,), (65536 + 3,), (1, 1), (3, 5), (9, 64), (130, 257),
                                   (1, 1, 1), (2, 3, 9), (8, 8, 8), (17, 33, 65), (40, 41, 130)])
def test_shapes_vs_oracle(gpu_ctx, shape):
    rng = np.random.default_rng(sum(shape))
    x = np.cumsum(rng.normal(0, 0.05, int(np.prod(shape)))).astype(np.float32).reshape(shape)
    for eb, r2r in ((1e-3, False), (1e-4, True)):
        if r2r and x.size == 1:
            continue  # zero value range: the reference exits, this library returns VECSZ_ERANGE0
        _check_vs_oracle(gpu_ctx, x, eb, r2r, tag=f"walk{shape}")


@pytest.mark.parametrize("block", [16, 32, 64])
def test_block_sizes_vs_oracle(gpu_ctx, block):
    from vecsz_amd import datagen as dg

    for shape in ((1000,), (70, 131), (33, 70, 75)):
        x = dg.grf(shape, seed=block) if len(shape) > 1 else np.cumsum(
            np.random.default_rng(block).normal(0, 0.1, shape)).astype(np.float32)
        _check_vs_oracle(gpu_ctx, x, 1e-3, True, block=block, tag=f"B{block}{shape}")


def test_dict_sizes_vs_oracle(gpu_ctx):
    rng = np.random.default_rng(12)
    x = rng.normal(0, 1, (30, 31, 33)).astype(np.float32)
    for d, eb in ((256, 1e-3), (1024, 1e-4), (65536, 1e-5), (2, 1e-1)):
        _check_vs_oracle(gpu_ctx, x, eb, False, dict_size=d, tag=f"dict{d}")


def test_special_values_vs_oracle(gpu_ctx):
    rng = np.random.default_rng(13)
    x = rng.normal(0, 1, (20, 21, 22)).astype(np.float32)
    x[0, 0, 0] = -0.0
    x[1, 2, 3] = np.inf
    x[4, 5, 20] = -np.inf
    x[7, 7, 7] = -1e-38
    x[8, 8, 8] = 1e-45
    _check_vs_oracle(gpu_ctx, x, 1e-2, False, tag="specials")
    big = (np.float32(3.3e7) + np.cumsum(rng.integers(-3, 4, 30 * 31 * 33))).astype(np.float32).reshape(30, 31, 33)
    _check_vs_oracle(gpu_ctx, big, 0.37, False, tag="beyond 2^24")
    _check_vs_oracle(gpu_ctx, big[0], 0.37, False, tag="beyond 2^24 2d")
    _check_vs_oracle(gpu_ctx, big.ravel()[:5000].copy(), 0.37, False, tag="beyond 2^24 1d")


@pytest.mark.parametrize("shape", [(300,), (20, 37), (9, 10, 21)])
def test_negative_zero_outliers(gpu_ctx, shape):
    """-0.0 right after a jump of >= R quanta is an outlier whose value is
    -0.0; the reference reconstructs it as -0.0."""
    x = np.zeros(shape, np.float32).ravel()
    x[5::7] = 5000.0
    x[6::7] = -0.0
    x = x.reshape(shape)
    _check_vs_oracle(gpu_ctx, x, 0.5, False, tag=f"-0 outliers {shape}")


@pytest.mark.parametrize("shape,blocks", [((24, 36, 68), (8, 16)), ((50, 264), (8, 16, 32, 64)),
                                          ((4100,), (8, 16, 32, 64))])
def test_four_per_lane_reconstruction(gpu_ctx, shape, blocks):
    """Rows that are multiples of 4 take the four-per-lane reconstruction;
    spikes put outliers (scan restarts) at every position inside a lane."""
    rng = np.random.default_rng(len(shape))
    x = np.cumsum(rng.normal(0, 0.05, int(np.prod(shape)))).astype(np.float32)
    x[rng.integers(0, x.size, x.size // 50)] += rng.choice([-1, 1], x.size // 50) * 500.0
    x = x.reshape(shape)
    for B in blocks:
        _check_vs_oracle(gpu_ctx, x, 1e-3, False, block=B, tag=f"x4 B{B}{shape}")
        _check_vs_oracle(gpu_ctx, x, 1e-4, True, block=B, tag=f"x4 r2r B{B}{shape}")


@pytest.mark.parametrize("shape,blocks,offset", [((9, 20, 131), (8, 16), 0), ((7, 13, 257), (8,), 0),
                                                 ((12, 10, 136), (8, 16), 1), ((40, 129), (8, 16, 32, 64), 0),
                                                 ((30, 256), (8, 64), 1)])
def test_unaligned_four_wide_tiles(gpu_ctx, shape, blocks, offset):
    """Fields at least 128 wide whose x extent is not a multiple of 4, or that
    start off a 16-byte boundary, take the 4-wide dual-quant tiles with
    element loads in a batch (a single field this small takes the 64-wide
    tiles); the partial x group (3D scalar forms, 2D lost columns) falls
    inside the last lane group.  Both forms against the oracle."""
    rng = np.random.default_rng(sum(shape) + offset)
    x = np.cumsum(rng.normal(0, 0.05, int(np.prod(shape)))).astype(np.float32)
    x[rng.integers(0, x.size, x.size // 100)] += 300.0
    x = x.reshape(shape)
    other = torch.from_numpy(np.cumsum(rng.normal(0, 0.05, 4000)).astype(np.float32)).cuda()
    for B in blocks:
        for eb, r2r in ((1e-3, False), (1e-4, True)):
            _check_vs_oracle(gpu_ctx, x, eb, r2r, block=B, tag=f"x4u B{B}{shape}", offset=offset)
            a_orc, n_orc, _ = po.compress(x, eb, r2r, 4096, B)
            buf = torch.zeros(x.size + offset, dtype=torch.float32, device="cuda")
            buf[offset:] = torch.from_numpy(x.ravel()).cuda()
            xt = buf[offset:].view(x.shape)
            mode = "r2r" if r2r else "abs"
            (a, nout), _ = gpu_ctx.compress_batch([xt, other], [eb, 1e-3], [mode, "abs"], block=B)
            assert a.cpu().numpy().tobytes() == a_orc and nout == n_orc, (shape, B, r2r, offset)


def test_c1_sinusoids_full_size(gpu_ctx):
    """2D 1800x3600 abs 1e-4, Huffman only: bit-exact archive."""
    from vecsz_amd import datagen as dg

    x = dg.sinusoids_2d(1800, 3600, seed=0)
    _check_vs_oracle(gpu_ctx, x, 1e-4, False, tag="C1")
    man = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
    d = {v["name"]: v for v in man["digests"]}["c1_sinusoids_2d_full"]
    if hashlib.sha256(x.tobytes()).hexdigest() == d["sha256_input"]:
        a, _ = gpu_ctx.compress(torch.from_numpy(x).cuda(), 1e-4, "abs")
        assert hashlib.sha256(a.cpu().numpy().tobytes()).hexdigest() == d["sha256_archive"]


def test_grf_256_r2r(gpu_ctx):
    from vecsz_amd import datagen as dg

    x = dg.grf((256, 256, 256), seed=2)
    for eb in (1e-2, 1e-4):
        _check_vs_oracle(gpu_ctx, x, eb, True, tag=f"grf256 {eb}")


def test_walk_spikes_1d_16m(gpu_ctx):
    from vecsz_amd import datagen as dg

    x = dg.walk_spikes_1d(1 << 24, seed=3)
    _check_vs_oracle(gpu_ctx, x, 1e-4, True, tag="walk1d 16M")


def test_repeat_is_deterministic(gpu_ctx):
    from vecsz_amd import datagen as dg

    x = torch.from_numpy(dg.grf((64, 96, 128), seed=9)).cuda()
    a1, _ = gpu_ctx.compress(x, 1e-4, "r2r")
    b1 = a1.cpu().numpy().tobytes()
    for _ in range(3):
        a2, _ = gpu_ctx.compress(x, 1e-4, "r2r")
        assert a2.cpu().numpy().tobytes() == b1


def test_errors_are_status_codes(gpu_ctx, golden):
    from vecsz_amd import VecszError

    c = [c for c in golden["cases"] if c["name"] == "3d_grf16_r2r_1e-2"][0]
    a = np.frombuffer(c["archive"], dtype=np.uint8).copy()
    # truncated archive
    with pytest.raises(VecszError):
        gpu_ctx.decompress(torch.from_numpy(a[: len(a) // 2].copy()).cuda(), "r2r", c["value_range"])
    # node count larger than the tree bytes present
    bad = a.copy()
    off = 4 * 4 + 8 + 24
    bad[off:off + 4] = np.frombuffer((10 ** 6).to_bytes(4, "big"), dtype=np.uint8)
    with pytest.raises(VecszError):
        gpu_ctx.decompress(torch.from_numpy(bad).cuda(), "r2r", c["value_range"])
    # r2r without a range, zero range
    with pytest.raises(VecszError):
        gpu_ctx.decompress(torch.from_numpy(a).cuda(), "r2r", 0.0)
    with pytest.raises(VecszError):
        gpu_ctx.compress(torch.zeros(64, 64, device="cuda"), 1e-3, "r2r")
    # output buffer too small
    with pytest.raises(VecszError):
        gpu_ctx.compress(torch.from_numpy(c["x"]).cuda(), c["eb"], "r2r", out=None, block=12)


def test_stage_timing(gpu_ctx):
    from vecsz_amd import datagen as dg

    x = torch.from_numpy(dg.grf((64, 64, 64), seed=1)).cuda()
    gpu_ctx.enable_timing(True)
    a, _ = gpu_ctx.compress(x, 1e-4, "r2r")
    names = [n for n, _ in gpu_ctx.timings()]
    assert "dualquant" in names and "pack" in names  # the three-kernel encode (the default)
    gpu_ctx.decompress(a, "r2r", float(x.max() - x.min()))
    assert "reconstruct" in [n for n, _ in gpu_ctx.timings()]
    gpu_ctx.enable_timing(False)
