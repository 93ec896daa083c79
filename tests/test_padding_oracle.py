"""Padding modes (--padding global|block|edge with max|min|avg|const) in the
C oracle: against the reference's own archives and outputs (golden fixtures
from the compiled reference), against the live reference build where it is
present, and against what the arithmetic fixes (closed forms, brute-force
pad values, equivalences between modes).  Runs without a GPU."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.golden.cases import pad_specified
from vecsz_amd import datagen as dg


def _sections_match(a: bytes, b: bytes, dims, block, mode, typ):
    sa, sb = po.archive_sections(a), po.archive_sections(b)
    if a[:sa["meta_bytes"]] != b[:sb["meta_bytes"]] or sa["huff"] != sb["huff"]:
        return False
    if not np.array_equal(sa["outliers"].view(np.uint32), sb["outliers"].view(np.uint32)):
        return False
    m = pad_specified(dims, block, mode, typ)
    pa, pb = sa["pad"].view(np.uint32), sb["pad"].view(np.uint32)
    return len(pa) == len(pb) == len(m) and np.array_equal(pa[m], pb[m])


def test_golden_padded_archives(golden):
    assert len(golden["pad_cases"]) >= 30
    bad = []
    for c in golden["pad_cases"]:
        a, nout, _ = po.compress(c["x"], c["eb"], c["r2r"], c["dict"], c["block"], c["pad_mode"], c["pad_type"],
                                 c["pad_constant"])
        dims = po.dims_xfirst(c["x"])
        if nout != c["n_outlier"] or not _sections_match(a, c["archive"], dims, c["block"], c["pad_mode"],
                                                         c["pad_type"]):
            bad.append(c["name"])
        # fully specified archives are byte-identical
        if c["pad_mode"] == "global" or c["pad_type"] == "const":
            if a != c["archive"]:
                bad.append(c["name"] + " (bytes)")
    assert not bad, f"oracle padded archive differs from the reference for {bad}"


def test_golden_padded_decompress(golden):
    bad = []
    for c in golden["pad_cases"]:
        y = po.decompress(c["archive"], c["x"].shape, c["r2r"], c["value_range"], c["dict"], c["block"])
        if not np.array_equal(y.view(np.uint32), c["out"].view(np.uint32)):
            bad.append(c["name"])
    assert not bad, f"oracle padded decompression differs from the reference for {bad}"


def test_golden_padded_covers_modes(golden):
    seen = {(c["pad_mode"], c["pad_type"], c["x"].ndim) for c in golden["pad_cases"]}
    for mode in ("global", "block", "edge"):
        for typ in ("max", "min", "avg", "const"):
            assert any((mode, typ, nd) in seen for nd in (1, 2, 3)), (mode, typ)
        for nd in (1, 2, 3):
            assert (mode, "const", nd) in seen
    assert any(c["block"] != 8 for c in golden["pad_cases"])


def _ebs(x, eb, r2r):
    return po.eb_family(r2r, eb, po.value_range(x) if r2r else 0.0)


@pytest.mark.parametrize("shape", [(64, 48), (16, 24, 40), (2048,)])
def test_pad_values_brute_force(shape):
    """Per-block max / min / avg against numpy on the block's elements; the
    2D/3D value is then prequantized as (float)((double)p * r) rounded half away."""
    x = dg.grf(shape, seed=21) if len(shape) > 1 else np.cumsum(
        np.random.default_rng(1).normal(0, 0.01, shape)).astype(np.float32)
    B, eb = 8, 1e-3
    r = _ebs(x, eb, False).r_d
    dims = po.dims_xfirst(x)
    for typ, fn in (("max", np.max), ("min", np.min), ("avg", None)):
        pv, halo = po.pad_values(x, eb, False, B, "block", typ)
        nb = [po.nblk(d, B) for d in dims]
        blocks = [(b0, b1, b2) for b2 in range(nb[2] if len(dims) > 2 else 1)
                  for b1 in range(nb[1] if len(dims) > 1 else 1) for b0 in range(nb[0])]
        assert len(pv) == len(blocks)
        for k, (b0, b1, b2) in enumerate(blocks):
            sl = [slice(b0 * B, b0 * B + B), slice(b1 * B, b1 * B + B), slice(b2 * B, b2 * B + B)][:len(dims)][::-1]
            blk = x[tuple(sl)]
            if blk.size == 0:
                assert pv[k] == 0.0
                continue
            if fn is not None:
                p = fn(blk)
            else:
                s = np.float32(0)
                for v in blk.ravel():
                    s = np.float32(s + v)
                p = np.float32(s / np.float32(blk.size))
            if len(dims) > 1:
                t = np.float32(np.float64(p) * r)
                p = np.float32(np.sign(t) * np.floor(np.abs(np.float64(t)) + 0.5))
            assert pv[k] == p, (typ, k)
            assert (halo[k] == pv[k]).all()


def test_pad_const_closed_form():
    x = dg.grf((24, 40), seed=2)
    eb, c = 1e-3, 3.3
    r = _ebs(x, eb, False).r_d
    pv, halo = po.pad_values(x, eb, False, 8, "global", "const", c)
    assert pv.size == 1 and pv[0] == np.float32(np.round(np.float64(np.float32(c)) * r))
    pv, halo = po.pad_values(x, eb, False, 8, "edge", "const", c)
    q = np.float32(np.round(np.float64(np.float32(np.float64(np.float32(c)) * r))))
    assert pv.size == 2 * po.nblk(40, 8) * po.nblk(24, 8) and (pv == q).all()


@pytest.mark.parametrize("mode", ["block", "edge", "global"])
def test_const_zero_padding_equals_no_padding(mode):
    """A pad value of +0.0 is the no-padding halo: the codes and outliers are
    those of an unpadded compress; only the pad section is added."""
    for shape in [(37, 45), (13, 21, 35), (1003,)]:
        x = dg.grf(shape, seed=4) if len(shape) > 1 else np.cumsum(
            np.random.default_rng(2).normal(0, 0.01, shape)).astype(np.float32)
        a0, n0, _ = po.compress(x, 1e-3, True)
        a1, n1, _ = po.compress(x, 1e-3, True, pad_mode=mode, pad_type="const", pad_constant=0.0)
        s0, s1 = po.archive_sections(a0), po.archive_sections(a1)
        assert n0 == n1 and s0["huff"] == s1["huff"]
        assert np.array_equal(s0["outliers"].view(np.uint32), s1["outliers"].view(np.uint32))
        assert (s1["pad"] == 0).all() and s1["pad_len"] > 0


@pytest.mark.parametrize("mode", ["global", "block", "edge"])
@pytest.mark.parametrize("typ", ["max", "min", "avg", "const"])
def test_padded_round_trip_error_bound(mode, typ):
    """Every padded mode keeps the error bound, except 1D block/edge padding:
    there the reference predicts a block's first element from the raw
    (not prequantized) pad value, so the prediction error is not integral and
    its conversion (rounding on the AVX path, truncation on the scalar one)
    can add up to one more quantization step, i.e. up to 3 eb in total."""
    for shape in [(40, 56), (16, 24, 32), (1024,)]:
        x = dg.grf(shape, seed=6) if len(shape) > 1 else dg.walk_spikes_1d(shape[0], seed=6)
        rng = po.value_range(x)
        a, nout, ebf = po.compress(x, 1e-3, True, pad_mode=mode, pad_type=typ, pad_constant=2.5)
        y = po.decompress(a, x.shape, True, rng)
        err = np.abs(y.astype(np.float64) - x.astype(np.float64)).max()
        k = 3.0 if (len(shape) == 1 and mode != "global") else 1.0
        assert err <= k * ebf * (1 + 1e-5) + 1e-6 * np.abs(x).max(), (shape, err, ebf)


def test_halo_from_archive_mode_detection():
    """pad_len 1 -> global, nblocks -> block, ndim * nblocks -> edge, else
    none (src/vecsz_interface.hh:287-298); 1D without a mode reads pad_vals[0]."""
    dims = (40, 24)
    nb = po.nblk(40, 8) * po.nblk(24, 8)
    pv = np.arange(1, 2 * nb + 1, dtype=np.float32)
    h = po.halo_from_archive(dims, 8, pv[:1])
    assert (h == 1).all()
    h = po.halo_from_archive(dims, 8, pv[:nb])
    assert (h[:, 0] == pv[:nb]).all() and (h[:, 2] == pv[:nb]).all()
    h = po.halo_from_archive(dims, 8, pv)
    assert (h[:, 2] == pv[0::2]).all() and (h[:, 1] == pv[1::2]).all()
    h = po.halo_from_archive(dims, 8, pv[:3])
    assert (h == 0).all()
    h1 = po.halo_from_archive((100,), 8, np.array([5.0, 6.0, 7.0], np.float32))
    assert (h1 == 5.0).all()


pytestmark_ref = pytest.mark.skipif(not po.ref_available(), reason="reference library not built here")


@pytestmark_ref
@pytest.mark.parametrize("shape", [(1000,), (40, 64), (37, 45), (16, 24, 32), (9, 10, 11)])
def test_oracle_vs_reference_padding(shape):
    rng = np.random.default_rng(sum(shape))
    x = np.cumsum(rng.normal(0, 0.05, int(np.prod(shape)))).astype(np.float32).reshape(shape)
    m = rng.random(shape) < 0.01
    x[m] += rng.uniform(5, 50, m.sum()).astype(np.float32)
    dims = po.dims_xfirst(x)
    full = all(d % 8 == 0 for d in dims)
    for mode in ("global", "block", "edge"):
        for typ in ("max", "min", "avg", "const"):
            if not full and mode != "global" and typ != "const":
                continue  # partial blocks: the reference reads uninitialized memory
            for r2r in (False, True):
                kw = dict(pad_mode=mode, pad_type=typ, pad_constant=-1.25)
                ra, rn = po.ref_compress(x, 1e-3, r2r, threads=1, **kw)
                oa, on, _ = po.compress(x, 1e-3, r2r, **kw)
                assert rn == on and _sections_match(oa, ra, dims, 8, mode, typ), (mode, typ, r2r)
                yr = po.ref_decompress(oa, x, r2r)
                yo = po.decompress(oa, x.shape, r2r, po.value_range(x) if r2r else 0.0).ravel()
                assert np.array_equal(yr.view(np.uint32), yo.view(np.uint32)), (mode, typ, r2r)
