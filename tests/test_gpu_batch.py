"""Batched compress/decompress (worker pool, one call for many fields):
every archive byte-identical to the single-field call and to the oracle,
every output bit-identical, mixed dimensionality, modes and error bounds."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _mix(seed=0):
    from vecsz_amd import datagen as dg

    rng = np.random.default_rng(seed)
    shapes = [(5000,), (1 << 16,), (3, 7), (64, 64), (130, 257), (512, 384), (8, 8, 8), (32, 32, 32),
              (17, 33, 65), (48, 64, 96), (64, 96, 128), (1,), (2, 3, 4), (1000, 4), (9, 20, 131)]
    fields, ebs, modes = [], [], []
    for i, s in enumerate(shapes):
        if len(s) >= 2 and min(s) >= 8:
            x = dg.grf(s, seed=seed + i)
        else:
            x = np.cumsum(rng.normal(0, 0.05, int(np.prod(s)))).astype(np.float32).reshape(s)
        fields.append(x)
        ebs.append(float(rng.choice([1e-2, 1e-3, 1e-4])))
        modes.append("abs" if (i % 3 == 0 or x.size == 1) else "r2r")
    return fields, ebs, modes


def test_batch_matches_single_calls_and_oracle(gpu_ctx):
    fields, ebs, modes = _mix()
    xs = [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in fields]
    print("batch of", len(xs), flush=True)
    res = gpu_ctx.compress_batch(xs, ebs, modes)
    for k, (x, f, eb, m) in enumerate(zip(xs, fields, ebs, modes)):
        a, nout = res[k]
        a1, n1 = gpu_ctx.compress(x, eb, m)
        ab = a.cpu().numpy().tobytes()
        assert ab == a1.cpu().numpy().tobytes() and nout == n1, k
        if f.size <= (1 << 17):
            a_orc, n_orc, _ = po.compress(f, eb, m == "r2r")
            assert ab == a_orc and nout == n_orc, k
    ranges = [po.value_range(f) if m == "r2r" else 0.0 for f, m in zip(fields, modes)]
    ys = gpu_ctx.decompress_batch([a for a, _ in res], modes, ranges)
    for k, (a, _) in enumerate(res):
        y1 = gpu_ctx.decompress(a, modes[k], ranges[k])
        assert torch.equal(ys[k].view(torch.int32), y1.view(torch.int32)), k
        if fields[k].size <= (1 << 17):
            y_orc = po.decompress(a.cpu().numpy().tobytes(), fields[k].shape, modes[k] == "r2r", ranges[k])
            assert np.array_equal(ys[k].cpu().numpy().view(np.uint32), y_orc.view(np.uint32)), k


def test_batch_is_deterministic_across_worker_counts(gpu_ctx):
    fields, ebs, modes = _mix(7)
    xs = [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in fields]
    ref = None
    for nw in (1, 3, 8):
        print("workers", nw, flush=True)
        gpu_ctx.set_workers(nw)
        out = [a.cpu().numpy().tobytes() for a, _ in gpu_ctx.compress_batch(xs, ebs, modes)]
        if ref is None:
            ref = out
        assert out == ref
    gpu_ctx.set_workers(8)


def test_batch_reports_a_failing_field(gpu_ctx):
    from vecsz_amd import VecszError

    xs = [torch.zeros(64, 64, device="cuda"), torch.randn(64, 64, device="cuda")]
    with pytest.raises(VecszError):
        gpu_ctx.compress_batch(xs, [1e-3, 1e-3], ["r2r", "r2r"])  # zero range in field 0
    assert len(gpu_ctx.compress_batch(xs, [1e-3, 1e-3], ["abs", "abs"])) == 2


def _many_small(seed=3, count=150):
    from vecsz_amd import datagen as dg

    rng = np.random.default_rng(seed)
    fields, ebs, modes = [], [], []
    for i in range(count):
        nd = int(rng.integers(1, 4))
        if nd == 1:
            s = (int(rng.integers(1, 20000)),)
        elif nd == 2:
            s = tuple(int(v) for v in rng.integers(1, 300, 2))
        else:
            s = tuple(int(v) for v in rng.integers(1, 60, 3))
        if min(s) >= 8 and nd >= 2:
            x = dg.grf(s, seed=seed + i)
        else:
            x = np.cumsum(rng.normal(0, 0.05, int(np.prod(s)))).astype(np.float32).reshape(s)
        fields.append(x)
        ebs.append(float(rng.choice([1e-2, 1e-3, 1e-4])))
        modes.append("abs" if (i % 4 == 0 or x.size == 1 or np.ptp(x) == 0) else "r2r")
    return fields, ebs, modes


@pytest.mark.parametrize("case", ["mix", "many_small", "block16", "dict1024"])
def test_fused_batch_matches_worker_path(gpu_ctx, case):
    """The fused compress batch (one launch per stage over all fields) and the
    worker-context path write the same archives."""
    block, dict_size = 8, 4096
    if case == "many_small":
        fields, ebs, modes = _many_small()
    else:
        fields, ebs, modes = _mix(11)
        if case == "block16":
            block = 16
        if case == "dict1024":
            dict_size = 1024
    xs = [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in fields]
    print("fused batch", case, len(xs), flush=True)
    try:
        fused = gpu_ctx.compress_batch(xs, ebs, modes, block=block, dict_size=dict_size)
        gpu_ctx.set_batch_fused(False)
        pool = gpu_ctx.compress_batch(xs, ebs, modes, block=block, dict_size=dict_size)
    finally:
        gpu_ctx.set_batch_fused(True)
    for k, ((a, n), (b, m)) in enumerate(zip(fused, pool)):
        assert n == m and a.cpu().numpy().tobytes() == b.cpu().numpy().tobytes(), (case, k, xs[k].shape)


def test_fused_batch_reports_a_small_output_buffer(gpu_ctx):
    from vecsz_amd import VecszError
    from vecsz_amd.interface import dims_of, make_params

    xs = [torch.randn(64, 64, device="cuda"), torch.randn(4096, device="cuda"), torch.randn(8, 8, 8, device="cuda")]
    params = [make_params(dims_of(x.shape), 1e-3, "abs", 8, 4096) for x in xs]
    caps = [gpu_ctx.compress_bound(p) for p in params]
    outs = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    a1, _ = gpu_ctx.compress(xs[1], 1e-3, "abs")
    caps[1] = a1.numel() - 1  # the header fits, the whole archive does not
    with pytest.raises(VecszError):
        gpu_ctx.compress_batch_into(params, [x.data_ptr() for x in xs], [o.data_ptr() for o in outs], caps)
    # the context stays usable and the next batch matches single calls
    res = gpu_ctx.compress_batch(xs, 1e-3, "abs")
    for x, (a, n) in zip(xs, res):
        a1, n1 = gpu_ctx.compress(x, 1e-3, "abs")
        assert a.cpu().numpy().tobytes() == a1.cpu().numpy().tobytes() and n == n1
