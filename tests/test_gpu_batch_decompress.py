"""Fused batch decompress (one launch per stage over all fields: the sync
pass, re-decode passes over one batch-wide list, one scan, the write pass,
reconstruction per tile form) against the worker-context path, single
calls and the oracle: mixed dimensionality, modes and error bounds, 150
small fields, B 16, dict 1024, one-symbol and all-outlier fields, outputs
that are not 16-byte aligned, and per-field failures (a corrupt header, a
short output buffer, a stream that holds fewer symbols than its header says)
that leave the other fields' outputs intact.  An archive with a pad section
sends the batch to the worker-context path."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from tests.test_gpu_batch import _many_small, _mix  # noqa: E402


def _archives(ctx, fields, ebs, modes, block=8, dict_size=4096):
    xs = [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in fields]
    return [a for a, _ in ctx.compress_batch(xs, ebs, modes, block=block, dict_size=dict_size)]


def _params(archives, modes, ranges, block, dict_size):
    from vecsz_amd import make_params

    return [make_params((1,), 1e-4, m, block, dict_size, r) for m, r in zip(modes, ranges)]


def _run(ctx, archives, params, outs, caps=None, check=True):
    caps = caps or [o.numel() for o in outs]
    return ctx.decompress_batch_into(params, [a.data_ptr() for a in archives], [a.numel() for a in archives],
                                     [o.data_ptr() for o in outs], caps, check=check)


@pytest.mark.parametrize("case", ["mix", "many_small", "block16", "dict1024", "special"])
def test_fused_decompress_batch_matches_worker_path_single_calls_and_oracle(gpu_ctx, case):
    block, dict_size = 8, 4096
    if case == "many_small":
        fields, ebs, modes = _many_small()
    elif case == "special":
        rng = np.random.default_rng(5)
        fields = [np.zeros((40, 50), np.float32),                                             # one symbol
                  (rng.standard_normal((9, 10, 11)) * 100).astype(np.float32),                # all outliers
                  np.cumsum(rng.normal(0, 0.05, 70001)).astype(np.float32),                   # 1D, d0 % 4 != 0
                  np.cumsum(rng.normal(0, 0.05, 4 * 9000)).astype(np.float32).reshape(90, 400)]
        ebs, modes = [1e-3, 1e-6, 1e-4, 1e-3], ["abs", "abs", "r2r", "r2r"]
    else:
        fields, ebs, modes = _mix(11)
        block = 16 if case == "block16" else 8
        dict_size = 1024 if case == "dict1024" else 4096
    print("fused decompress batch", case, len(fields), flush=True)
    archives = _archives(gpu_ctx, fields, ebs, modes, block, dict_size)
    ranges = [po.value_range(f) if m == "r2r" else 0.0 for f, m in zip(fields, modes)]
    params = _params(archives, modes, ranges, block, dict_size)
    fused = [torch.empty(f.shape, dtype=torch.float32, device="cuda") for f in fields]
    _run(gpu_ctx, archives, params, fused)  # small fields: the fused path by default
    try:
        gpu_ctx.set_batch_fused(False)
        pool = [torch.empty(f.shape, dtype=torch.float32, device="cuda") for f in fields]
        _run(gpu_ctx, archives, params, pool)
    finally:
        gpu_ctx.set_batch_fused(True)
    for k, a in enumerate(archives):
        y1 = gpu_ctx.decompress(a, modes[k], ranges[k], block=block, dict_size=dict_size)
        assert torch.equal(fused[k].view(torch.int32), y1.view(torch.int32)), (case, k)
        assert torch.equal(pool[k].view(torch.int32), y1.view(torch.int32)), (case, k)
        if fields[k].size <= (1 << 17):
            y_orc = po.decompress(a.cpu().numpy().tobytes(), fields[k].shape, modes[k] == "r2r", ranges[k],
                                  dict_size, block)
            assert np.array_equal(fused[k].cpu().numpy().view(np.uint32), y_orc.view(np.uint32)), (case, k)


def test_fused_decompress_batch_of_slowly_synchronising_walks(gpu_ctx):
    """1D random walks with spikes synchronise slowly: their re-decode lists
    hold chains (a subsequence's corrected end moves its successor's start,
    and so on) that take several passes over the batch-wide list.  Large
    fields fused ("always"), twice, against single calls and the oracle."""
    from vecsz_amd import datagen as dg

    fields = [dg.walk_spikes_1d(n, seed=s) for s, n in ((1, 1 << 22), (2, 3 << 20), (3, (1 << 21) + 77))]
    fields.append(dg.grf((300, 500), seed=4))
    ebs, modes = [1e-4, 1e-4, 1e-5, 1e-4], ["r2r", "r2r", "r2r", "abs"]
    print("fused decompress batch of walks", [f.size for f in fields], flush=True)
    archives = _archives(gpu_ctx, fields, ebs, modes)
    ranges = [po.value_range(f) if m == "r2r" else 0.0 for f, m in zip(fields, modes)]
    params = _params(archives, modes, ranges, 8, 4096)
    fused = [torch.empty(f.shape, dtype=torch.float32, device="cuda") for f in fields]
    try:
        gpu_ctx.set_batch_fused("always")
        for rep in range(2):
            for o in fused:
                o.fill_(-1.0)
            _run(gpu_ctx, archives, params, fused)
            for k, a in enumerate(archives):
                y1 = gpu_ctx.decompress(a, modes[k], ranges[k])
                assert torch.equal(fused[k].view(torch.int32), y1.view(torch.int32)), (rep, k)
    finally:
        gpu_ctx.set_batch_fused(True)
    for k, a in enumerate(archives):
        y_orc = po.decompress(a.cpu().numpy().tobytes(), fields[k].shape, modes[k] == "r2r", ranges[k], 4096, 8)
        assert np.array_equal(fused[k].cpu().numpy().view(np.uint32), y_orc.view(np.uint32)), k


def test_fused_decompress_batch_unaligned_outputs(gpu_ctx):
    """Outputs one float past a 16-byte boundary take the element-wise
    reconstruction forms; the values are the same."""
    fields, ebs, modes = _mix(13)
    archives = _archives(gpu_ctx, fields, ebs, modes)
    ranges = [po.value_range(f) if m == "r2r" else 0.0 for f, m in zip(fields, modes)]
    params = _params(archives, modes, ranges, 8, 4096)
    bufs = [torch.empty(f.size + 1, dtype=torch.float32, device="cuda") for f in fields]
    outs = [b[1:] for b in bufs]
    _run(gpu_ctx, archives, params, outs)
    for k, a in enumerate(archives):
        y1 = gpu_ctx.decompress(a, modes[k], ranges[k])
        assert torch.equal(outs[k].view(torch.int32), y1.reshape(-1).view(torch.int32)), k


def test_fused_decompress_batch_reports_failing_fields(gpu_ctx):
    from vecsz_amd._lib import VECSZ_ECORRUPT, VECSZ_ENOMEM

    fields, ebs, modes = _mix(17)
    fields, ebs, modes = fields[:8], ebs[:8], ["abs"] * 8
    archives = _archives(gpu_ctx, fields, ebs, modes)
    ranges = [0.0] * 8
    # field 3: corrupt header (ndim 7)
    archives[3] = archives[3].clone()
    archives[3][0] = 7
    # field 1 (1D, 65536 elements): its header claims twice the elements, so
    # the stream holds fewer symbols than the field
    assert fields[1].shape == (1 << 16,)
    archives[1] = archives[1].clone()
    d = np.frombuffer(archives[1][4:8].cpu().numpy().tobytes(), np.int32)[0]
    archives[1][4:8] = torch.from_numpy(np.array([2 * d], np.int32).view(np.uint8)).cuda()
    params = _params(archives, modes, ranges, 8, 4096)
    outs = [torch.full((max(f.size, 1) * 2,), -1.0, device="cuda") for f in fields]
    caps = [f.size for f in fields]
    caps[1] = 2 * fields[1].size
    caps[6] = fields[6].size - 1  # field 6: output one element short
    for nw in (True, False):  # the fused path, then the worker contexts
        gpu_ctx.set_batch_fused(nw)
        try:
            for o in outs:
                o.fill_(-1.0)
            rc, st = _run(gpu_ctx, archives, params, outs, caps, check=False)
        finally:
            gpu_ctx.set_batch_fused(True)
        assert rc != 0
        assert st[1] == VECSZ_ECORRUPT and st[3] == VECSZ_ECORRUPT and st[6] == VECSZ_ENOMEM, (nw, st)
        for k in range(8):
            if k in (1, 3, 6):
                continue
            assert st[k] == 0, (nw, k, st)
            y1 = gpu_ctx.decompress(archives[k], modes[k], ranges[k])
            assert torch.equal(outs[k][:fields[k].size].view(torch.int32), y1.reshape(-1).view(torch.int32)), (nw, k)


def test_padded_archive_takes_the_worker_path(gpu_ctx):
    from vecsz_amd import datagen as dg

    xs = [torch.from_numpy(dg.grf((40, 48, 56), seed=s)).cuda() for s in (1, 2, 3)]
    archives = [gpu_ctx.compress(xs[0], 1e-3, "abs")[0],
                gpu_ctx.compress(xs[1], 1e-3, "abs", pad_mode="block", pad_type="avg")[0],
                gpu_ctx.compress(xs[2], 1e-3, "abs")[0]]
    params = _params(archives, ["abs"] * 3, [0.0] * 3, 8, 4096)
    outs = [torch.empty_like(x) for x in xs]
    _run(gpu_ctx, archives, params, outs)
    for k in range(3):
        y1 = gpu_ctx.decompress(archives[k], "abs")
        assert torch.equal(outs[k].view(torch.int32), y1.view(torch.int32)), k


def test_batch_of_large_fields_takes_the_worker_path_unless_asked(gpu_ctx):
    """Fields of more than 4M elements on average decompress on the worker
    contexts by default; "always" runs them fused; the outputs agree."""
    from vecsz_amd import datagen as dg

    xs = [dg.grf((64, 256, 320), seed=31, device="cuda", pow2=True),
          dg.grf((80, 256, 256), seed=32, device="cuda", pow2=True)]  # 5.2M elements each
    archives = [gpu_ctx.compress(x, 1e-3, "r2r")[0] for x in xs]
    ranges = [float((x.max() - x.min()).item()) for x in xs]
    params = _params(archives, ["r2r"] * 2, ranges, 8, 4096)
    outs = {}
    for mode in (True, "always"):
        gpu_ctx.set_batch_fused(mode)
        try:
            outs[mode] = [torch.empty_like(x) for x in xs]
            _run(gpu_ctx, archives, params, outs[mode])
        finally:
            gpu_ctx.set_batch_fused(True)
    for k in range(2):
        y1 = gpu_ctx.decompress(archives[k], "r2r", ranges[k])
        for mode in outs:
            assert torch.equal(outs[mode][k].view(torch.int32), y1.view(torch.int32)), (mode, k)


def test_set_batch_fused_modes(gpu_ctx):
    from vecsz_amd import VecszError

    for mode in (False, True, "always", True):
        gpu_ctx.set_batch_fused(mode)
    rc = gpu_ctx._lib.vecsz_ctx_set_batch_fused(gpu_ctx._ctx, 3)
    assert rc != 0
    gpu_ctx.set_batch_fused(True)
    with pytest.raises(VecszError):
        gpu_ctx._check(rc)
