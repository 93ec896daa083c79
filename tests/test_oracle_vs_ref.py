"""The C oracle against the compiled reference on fresh random inputs (runs
where oracle/_ref/libvecsz_ref.so was built, i.e. where the reference
sources are present)."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not po.ref_available(), reason="reference library not built here")


def _field(rng, shape, kind):
    n = int(np.prod(shape))
    if kind == "walk":
        return np.cumsum(rng.normal(0, 0.01, n)).astype(np.float32).reshape(shape)
    if kind == "noise":
        return rng.normal(0, 1, shape).astype(np.float32)
    if kind == "spiky":
        x = np.cumsum(rng.normal(0, 0.01, n)).astype(np.float32).reshape(shape)
        m = rng.random(shape) < 0.02
        x[m] += (rng.uniform(10, 1000, m.sum()) * rng.choice([-1, 1], m.sum())).astype(np.float32)
        return x
    g = np.meshgrid(*[np.linspace(0, 3, s) for s in shape], indexing="ij")
    return (np.sin(sum(g)) + rng.normal(0, 1e-3, shape)).astype(np.float32)


SHAPES = [(100,), (1000,), (37, 45), (24, 33), (9, 10, 11), (16, 16, 16), (13, 20, 27), (8, 8, 17)]


@pytest.mark.parametrize("shape", SHAPES)
def test_random_vs_reference(shape):
    rng = np.random.default_rng(hash(shape) & 0xFFFF)
    for kind in ("walk", "noise", "smooth", "spiky"):
        for eb, r2r in ((1e-3, False), (1e-2, True), (1e-4, True)):
            for block in (8, 16):
                x = _field(rng, shape, kind)
                a_ref, n_ref = po.ref_compress(x, eb, r2r, block=block)
                a_orc, n_orc, _ = po.compress(x, eb, r2r, block=block)
                assert a_ref == a_orc and n_ref == n_orc, (shape, kind, eb, r2r, block)
                rng_v = po.value_range(x) if r2r else 0.0
                y_ref = po.ref_decompress(a_ref, x, r2r, block=block, pad0=(x.ndim == 1)).reshape(x.shape)
                y_orc = po.decompress(a_orc, x.shape, r2r, rng_v, block=block)
                assert np.array_equal(y_ref.view(np.uint32), y_orc.view(np.uint32)), (shape, kind, eb, r2r, block)


def test_reference_2d_decompress_matches_pad0_variant():
    """The pad-0 decompress driver used for 1D fixtures reproduces the
    reference's own Decompress where the latter is well defined (2D/3D)."""
    rng = np.random.default_rng(5)
    for shape in ((30, 41), (9, 12, 17)):
        x = _field(rng, shape, "smooth")
        a, _ = po.ref_compress(x, 1e-3)
        y1 = po.ref_decompress(a, x)
        y2 = po.ref_decompress(a, x, pad0=True)
        assert np.array_equal(y1.view(np.uint32), y2.view(np.uint32))
