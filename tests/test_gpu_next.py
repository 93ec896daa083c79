"""The rows after the core path: metadata sidecar, on-device verification
(VerifyData), and the lossless pass through the reference-shaped interface."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _verify_ref(o: np.ndarray, x: np.ndarray) -> dict:
    """VerifyData (src/utils/verify.hh:30-74) restated with numpy: float32
    per-element errors, float64 sums."""
    o = o.ravel().astype(np.float32)
    x = x.ravel().astype(np.float32)
    n = o.size
    ae = np.abs(x - o)
    nz = o != 0
    rel = (ae[nz] / np.abs(o[nz])).max() if nz.any() else 0.0
    mo, mx_ = o.astype(np.float64).mean(), x.astype(np.float64).mean()
    do, dx = o.astype(np.float64) - mo, x.astype(np.float64) - mx_
    mse = (ae * ae).astype(np.float64).sum() / n
    rng = float(o.max()) - float(o.min())
    return {"minimum": float(o.min()), "maximum": float(o.max()), "range": rng, "max_abserr": float(ae.max()),
            "max_abserr_index": int(np.argmax(ae)), "max_pwr_rel_abserr": float(rel), "MSE": mse,
            "NRMSE": np.sqrt(mse) / rng, "PSNR": 20 * np.log10(rng) - 10 * np.log10(mse),
            "coeff": (do * dx).sum() / n / np.sqrt((do * do).sum() / n) / np.sqrt((dx * dx).sum() / n)}


def test_sidecar_makes_decompress_self_sufficient(gpu_ctx):
    from vecsz_amd import datagen as dg, make_params, params_from_sidecar

    x = torch.from_numpy(dg.grf((40, 48, 64), seed=4)).cuda()
    for mode, eb in (("r2r", 1e-4), ("abs", 0.05)):
        a, _ = gpu_ctx.compress(x, eb, mode, block=16, dict_size=1024)
        rec = gpu_ctx.last_sidecar()
        assert len(rec) == 48
        p = params_from_sidecar(rec)
        rng = float((x.max() - x.min()).item()) if mode == "r2r" else 0.0
        assert p.mode == (1 if mode == "r2r" else 0) and p.block == 16 and p.dict_size == 1024
        assert p.value_range == rng
        y1 = torch.empty_like(x)
        gpu_ctx.decompress_into(p, a.data_ptr(), a.numel(), y1.data_ptr(), y1.numel())
        y2 = gpu_ctx.decompress(a, mode, rng, 16, 1024)
        assert torch.equal(y1.view(torch.int32), y2.view(torch.int32)), mode


def test_verify_matches_verifydata(gpu_ctx):
    from vecsz_amd import datagen as dg

    x = dg.grf((30, 40, 50), seed=6)
    for eb in (1e-2, 1e-4):
        a, _ = gpu_ctx.compress(torch.from_numpy(x).cuda(), eb, "r2r")
        y = gpu_ctx.decompress(a, "r2r", po.value_range(x))
        st = gpu_ctx.verify(torch.from_numpy(x).cuda(), y)
        ref = _verify_ref(x, y.cpu().numpy())
        for k in ("minimum", "maximum", "range", "max_abserr", "max_abserr_index", "max_pwr_rel_abserr"):
            assert st[k] == ref[k], (k, st[k], ref[k])
        for k in ("MSE", "NRMSE", "PSNR", "coeff"):
            assert st[k] == pytest.approx(ref[k], rel=1e-9), k
        assert st["len"] == x.size
        assert st["max_abserr"] <= eb * po.value_range(x) * (1 + 1e-5)


def test_verify_first_index_of_the_maximum(gpu_ctx):
    o = torch.zeros(1 << 20, device="cuda")
    x = torch.zeros(1 << 20, device="cuda")
    o[5] = 1.0
    for i in (700_000, 1234, 999_999):
        x[i] = 0.5
    st = gpu_ctx.verify(o, x)
    assert st["max_abserr"] == 1.0 and st["max_abserr_index"] == 5
    x[5] = 1.0  # error at 5 is now 0: ties at 0.5, the first index wins
    st = gpu_ctx.verify(o, x)
    assert st["max_abserr"] == 0.5 and st["max_abserr_index"] == 1234


@pytest.mark.parametrize("kind", ["gzip", "zstd"])
def test_reference_shaped_interface_with_lossless_pass(golden, kind):
    from vecsz_amd import ArgParse, Compress, Decompress, Dim4, lossless_compress

    c = [c for c in golden["cases"] if c["name"] == "3d_grf_partial_x_1e-4"][0]
    ap = ArgParse(ndim=3, dim4=Dim4(*reversed(c["x"].shape)), eb=c["eb"], mode="r2r", value_range=c["value_range"],
                  lossless=kind)
    stream, nout, size, lsize = Compress(ap, c["x"].ravel())
    assert size == len(c["archive"]) and lsize == len(stream) and lsize > 0
    if kind == "gzip":
        assert stream == lossless_compress("gzip", c["archive"])
    ap2 = ArgParse(mode="r2r", value_range=c["value_range"])
    y = Decompress(ap2, stream, size, lsize)
    assert np.array_equal(np.asarray(y).view(np.uint32), c["out"].view(np.uint32))
