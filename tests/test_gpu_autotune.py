"""Launch-shape autotuner (vecsz_autotune): the shapes it keeps on the context
leave every archive unchanged, single and batched, and the per-block-size
report matches plain compresses with that block size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_autotune_keeps_archives_identical():
    from vecsz_amd import Context, datagen

    ctx = Context(0)
    x = datagen.grf((64, 96, 128), seed=21, device="cuda", pow2=True)
    y = datagen.grf((300, 257), seed=22, device="cuda", pow2=True)
    a0, n0 = ctx.compress(x, 1e-3, "r2r")
    b0, m0 = ctx.compress(y, 1e-4, "abs")
    print("autotune 64x96x128", flush=True)
    t = ctx.autotune(x, 1e-3, "r2r")
    assert t["block"] in (8, 16, 32, 64)
    assert t["dq_grid"] in (256, 512, 768, 1024, 1536, 2048, 4096)
    assert t["enc_grid"] in (256, 512, 768, 1024, 2048, 4096, 8192)
    assert t["dq_ms"] > 0 and t["enc_ms"] > 0
    assert t["bytes_by_block"][8] == a0.numel()
    a16, _ = ctx.compress(x, 1e-3, "r2r", block=16)
    assert t["bytes_by_block"][16] == a16.numel()
    a1, n1 = ctx.compress(x, 1e-3, "r2r")
    b1, m1 = ctx.compress(y, 1e-4, "abs")
    assert a1.cpu().numpy().tobytes() == a0.cpu().numpy().tobytes() and n1 == n0
    assert b1.cpu().numpy().tobytes() == b0.cpu().numpy().tobytes() and m1 == m0
    # batched forms on the tuned context
    for fused in (True, False):
        ctx.set_batch_fused(fused)
        res = ctx.compress_batch([x, y], [1e-3, 1e-4], ["r2r", "abs"])
        assert res[0][0].cpu().numpy().tobytes() == a0.cpu().numpy().tobytes()
        assert res[1][0].cpu().numpy().tobytes() == b0.cpu().numpy().tobytes()
    ctx.set_batch_fused(True)
