"""Padding modes on the HIP path (through the C ABI): against the reference's
padded golden archives and outputs, and against the C oracle on fresh
inputs (every mode x value type x dimensionality, partial and trailing
blocks, every block size), byte for byte and bit for bit."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.golden.cases import pad_specified
from vecsz_amd import datagen as dg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _mode(r2r):
    return "r2r" if r2r else "abs"


def _gpu_compress(ctx, x, eb, r2r, block, mode, typ, const):
    a, nout = ctx.compress(torch.from_numpy(np.ascontiguousarray(x)).cuda(), eb, _mode(r2r), block, 4096,
                           pad_mode=mode, pad_type=typ, pad_constant=const)
    return a.cpu().numpy().tobytes(), nout


def _gpu_decompress(ctx, a, shape, r2r, rng, block):
    t = torch.from_numpy(np.frombuffer(a, dtype=np.uint8).copy()).cuda()
    return ctx.decompress(t, _mode(r2r), rng, block).cpu().numpy().reshape(shape)


def test_golden_padded_compress(gpu_ctx, golden):
    bad = []
    for c in golden["pad_cases"]:
        _say("compress", c["name"])
        a, nout = _gpu_compress(gpu_ctx, c["x"], c["eb"], c["r2r"], c["block"], c["pad_mode"], c["pad_type"],
                                c["pad_constant"])
        sa, sr = po.archive_sections(a), po.archive_sections(c["archive"])
        m = pad_specified(po.dims_xfirst(c["x"]), c["block"], c["pad_mode"], c["pad_type"])
        ok = (nout == c["n_outlier"] and a[:sa["meta_bytes"]] == c["archive"][:sr["meta_bytes"]]
              and sa["huff"] == sr["huff"]
              and np.array_equal(sa["outliers"].view(np.uint32), sr["outliers"].view(np.uint32))
              and len(sa["pad"]) == len(m)
              and np.array_equal(sa["pad"].view(np.uint32)[m], sr["pad"].view(np.uint32)[m]))
        # and byte for byte with the oracle, whose unspecified pad values are the +0.0 model's
        oa, _, _ = po.compress(c["x"], c["eb"], c["r2r"], c["dict"], c["block"], c["pad_mode"], c["pad_type"],
                               c["pad_constant"])
        if not ok or a != oa:
            bad.append(c["name"])
    assert not bad, bad


def test_golden_padded_decompress(gpu_ctx, golden):
    bad = []
    for c in golden["pad_cases"]:
        _say("decompress", c["name"])
        y = _gpu_decompress(gpu_ctx, c["archive"], c["x"].shape, c["r2r"], c["value_range"], c["block"])
        if not np.array_equal(y.view(np.uint32), c["out"].view(np.uint32)):
            bad.append(c["name"])
    assert not bad, bad


SHAPES = [(1000,), (4096,), (777,), (37, 45), (64, 96), (33, 130), (13, 21, 35), (16, 24, 32), (9, 40, 66)]


@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
def test_padded_vs_oracle(gpu_ctx, shape):
    rng = np.random.default_rng(sum(shape))
    x = dg.grf(shape, seed=len(shape) + shape[0]) if len(shape) > 1 else dg.walk_spikes_1d(shape[0], seed=7)
    m = rng.random(shape) < 0.005
    x[m] += rng.uniform(10, 100, m.sum()).astype(np.float32)
    bad = []
    for block in (8, 16, 32, 64):
        for mode in ("global", "block", "edge"):
            for typ in ("max", "min", "avg", "const"):
                r2r = (block + len(mode)) % 2 == 0
                eb = 1e-3 if block != 32 else 1e-4
                rngv = po.value_range(x) if r2r else 0.0
                name = f"{shape} B{block} {mode} {typ} {'r2r' if r2r else 'abs'}"
                _say(name)
                a, nout = _gpu_compress(gpu_ctx, x, eb, r2r, block, mode, typ, -2.5)
                oa, on, _ = po.compress(x, eb, r2r, 4096, block, mode, typ, -2.5)
                if a != oa or nout != on:
                    bad.append(name + " archive")
                    continue
                y = _gpu_decompress(gpu_ctx, a, x.shape, r2r, rngv, block)
                yo = po.decompress(oa, x.shape, r2r, rngv, 4096, block)
                if not np.array_equal(y.view(np.uint32), yo.view(np.uint32)):
                    bad.append(name + " decompress")
    assert not bad, bad


def test_padded_global_extremes(gpu_ctx):
    """Global max / min keep the first of equal values (the sign of a zero
    extreme) and a leading NaN; avg is the float running sum in index order."""
    cases = []
    a = -np.abs(dg.grf((16, 20, 24), seed=3))
    a.ravel()[5] = -0.0
    a.ravel()[900] = 0.0
    cases.append(("zero_max_first_neg", a))
    b = np.abs(dg.grf((16, 20, 24), seed=4))
    b.ravel()[7] = 0.0
    b.ravel()[70] = -0.0
    cases.append(("zero_min_first_pos", b))
    c = dg.grf((16, 20, 24), seed=5)
    c.ravel()[0] = np.nan
    cases.append(("leading_nan", c))
    d = dg.grf((16, 20, 24), seed=6)
    d.ravel()[100] = np.nan
    cases.append(("inner_nan", d))
    bad = []
    for name, x in cases:
        for typ in ("max", "min", "avg"):
            _say(name, typ)
            ga, gn = _gpu_compress(gpu_ctx, x, 1e-2, False, 8, "global", typ, 0.0)
            oa, on, _ = po.compress(x, 1e-2, False, 4096, 8, "global", typ, 0.0)
            if ga != oa or gn != on:
                bad.append((name, typ))
    assert not bad, bad


def test_padded_global_avg_large(gpu_ctx):
    """find_avg over 2^22 elements: the one-wavefront running sum matches the
    oracle's sequential float sum bit for bit."""
    x = dg.grf((64, 256, 256), seed=9)
    pv, _ = po.pad_values(x, 1e-3, False, 8, "global", "avg")
    a, _ = _gpu_compress(gpu_ctx, x, 1e-3, False, 8, "global", "avg", 0.0)
    assert po.archive_sections(a)["pad"].view(np.uint32)[0] == pv.view(np.uint32)[0]


@pytest.mark.parametrize("mode", ["block", "edge", "global"])
def test_padded_full_size_round_trip(gpu_ctx, mode):
    """A 256^3 field: archive equal to the oracle's, error bound kept."""
    x = dg.grf((256, 256, 256), seed=1)
    rngv = po.value_range(x)
    a, nout = _gpu_compress(gpu_ctx, x, 1e-4, True, 8, mode, "max", 0.0)
    oa, on, ebf = po.compress(x, 1e-4, True, 4096, 8, mode, "max", 0.0)
    assert a == oa and nout == on
    y = _gpu_decompress(gpu_ctx, a, x.shape, True, rngv, 8)
    err = np.abs(y.astype(np.float64) - x.astype(np.float64)).max()
    assert err <= ebf * (1 + 1e-5) + 2.0 ** -22 * np.abs(x).max()
