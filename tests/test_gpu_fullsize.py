"""Parity at the measured sizes: every archive byte and every decompressed
float of the HIP path equal the oracle's (the C restatement pinned to the
reference by the golden vectors and the live comparisons), at

  * 3D 512^3 GRF, r2r eb 1e-2, 1e-3 and 1e-4 (the 512 MiB configuration),
  * 1D 2^28 random walk with 0.1% spikes, r2r 1e-4 (1 GiB),
  * 3D 512x512x1024 GRF (x first), r2r 1e-4 (the 1 GiB target field),

for the archive-only decode and the decode-index path, and for the ~8 GiB
96-field batched mix through one fused batch call (src/vecsz_interface.hh:31-229
and :232-375 on every field).  The fields are generated on the GPU; the
oracle runs on host threads beside the GPU work.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

_POOL_THREADS = int(os.environ.get("VECSZ_TEST_THREADS", "16"))


def _say(*a):
    print(*a, flush=True)


def _oracle(x_host, eb, r2r):
    """Oracle archive, outlier count and decompressed field of one host array."""
    a, nout, _ = po.compress(x_host, eb, r2r)
    rng = po.value_range(x_host) if r2r else 0.0
    y = po.decompress(a, x_host.shape, r2r, rng)
    return a, nout, y, rng


def _check_field(ctx, pool, x_dev, eb, name):
    x = x_dev.cpu().numpy()
    _say(f"{name}: oracle started", tuple(x.shape), eb)
    fut = pool.submit(_oracle, x, eb, True)
    a, nout = ctx.compress(x_dev, eb, "r2r")
    a2, nout2, idx = ctx.compress_indexed(x_dev, eb, "r2r")
    assert nout2 == nout and torch.equal(a2, a)
    rng = float((x_dev.max() - x_dev.min()).item())
    y = ctx.decompress(a, "r2r", rng)
    y2 = ctx.decompress(a, "r2r", rng, index=idx)
    a_host = a.cpu().numpy().tobytes()
    y_host = y.cpu().numpy().reshape(-1)
    assert torch.equal(y2.view(torch.int32), y.view(torch.int32)), f"{name}: indexed decode differs"
    del y, y2
    want, nout_ref, y_ref, rng_ref = fut.result()
    _say(f"{name}: {len(a_host)} bytes, {nout} outliers")
    assert rng == rng_ref
    assert nout == nout_ref
    assert len(a_host) == len(want), f"{name}: archive size {len(a_host)} vs {len(want)}"
    assert a_host == want, f"{name}: archive differs at byte " + str(
        int(np.argmax(np.frombuffer(a_host, np.uint8) != np.frombuffer(want, np.uint8))))
    assert np.array_equal(y_host.view(np.uint32), y_ref.reshape(-1).view(np.uint32)), f"{name}: output differs"


@pytest.fixture(scope="module")
def pool():
    with cf.ThreadPoolExecutor(max_workers=_POOL_THREADS) as ex:
        yield ex


@pytest.mark.parametrize("eb", [1e-2, 1e-3, 1e-4])
def test_grf_512_cubed(gpu_ctx, pool, eb):
    from vecsz_amd import datagen as dg

    x = dg.grf((512, 512, 512), seed=0, device="cuda")
    _check_field(gpu_ctx, pool, x, eb, "512^3 GRF")


def test_walk_2p28(gpu_ctx, pool):
    from vecsz_amd import datagen as dg

    x = dg.walk_spikes_1d(1 << 28, seed=0, device="cuda")
    _check_field(gpu_ctx, pool, x, 1e-4, "2^28 walk")


def test_grf_1gib_target(gpu_ctx, pool):
    from vecsz_amd import datagen as dg

    x = dg.grf((1024, 512, 512), seed=0, device="cuda")  # the bench's target field (x 512, y 512, z 1024)
    _check_field(gpu_ctx, pool, x, 1e-4, "512x512x1024 GRF")


def test_batch_mix_8gib(gpu_ctx, pool):
    """The ~8 GiB 96-field mix (datagen.batch_mix) in one fused batch
    compress and one batch decompress; every field against the oracle."""
    from vecsz_amd import datagen as dg, make_params

    mix, seed = dg.batch_mix(96, seed=0)
    fields, params = [], []
    for k, (shape, eb) in enumerate(mix):
        fields.append(dg.grf(shape, seed=100 + k, device="cuda", pow2=True))
        params.append(make_params(tuple(reversed(shape)), eb, "r2r"))
    torch.cuda.synchronize()
    total = sum(4 * f.numel() for f in fields)
    _say(f"batch mix: {len(fields)} fields, {total / 2**30:.3f} GiB (draw seed {seed})")
    futs = [pool.submit(_oracle, f.cpu().numpy(), eb, True) for f, (_, eb) in zip(fields, mix)]
    caps = [gpu_ctx.compress_bound(p) for p in params]
    outs = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    res = gpu_ctx.compress_batch_into(params, [f.data_ptr() for f in fields], [o.data_ptr() for o in outs], caps)
    rngs = [float((f.max() - f.min()).item()) for f in fields]
    dparams = [make_params((1,), 1e-4, "r2r", value_range=r) for r in rngs]
    ys = [torch.empty_like(f) for f in fields]
    gpu_ctx.decompress_batch_into(dparams, [o.data_ptr() for o in outs], [s for s, _ in res],
                                  [y.data_ptr() for y in ys], [y.numel() for y in ys])
    torch.cuda.synchronize()
    for k, (fut, (size, nout)) in enumerate(zip(futs, res)):
        want, nout_ref, y_ref, _ = fut.result()
        assert nout == nout_ref, k
        assert size == len(want) and outs[k][:size].cpu().numpy().tobytes() == want, f"field {k} {mix[k]}"
        assert np.array_equal(ys[k].cpu().numpy().reshape(-1).view(np.uint32), y_ref.reshape(-1).view(np.uint32)), \
            f"field {k} output"
        futs[k] = None
    _say("batch mix: all 96 archives and outputs equal the oracle's")
