"""The request server: compress and decompress requests submitted together
(overlapping on the GPU) give exactly what each gives alone; admission keeps
to the HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _requests(ctx, seed=0):
    from vecsz_amd import datagen as dg, make_params

    rng = np.random.default_rng(seed)
    shapes = [(100_000,), (3000,), (96, 130), (256, 256), (20, 33, 70), (64, 64, 64), (1 << 18,), (40, 48, 56)]
    reqs = []
    for i, s in enumerate(shapes):
        x = torch.from_numpy(dg.grf(s, seed=i) if len(s) > 1 else
                             np.cumsum(rng.normal(0, 0.05, s)).astype(np.float32)).cuda()
        eb = float(rng.choice([1e-2, 1e-3, 1e-4]))
        p = make_params(tuple(reversed(s)), eb, "abs")
        a, nout = ctx.compress(x, eb, "abs")  # alone
        y = ctx.decompress(a, "abs")
        reqs.append({"x": x, "p": p, "a": a.clone(), "nout": nout, "y": y.clone()})
    return reqs


def _run_mixed(server, reqs):
    from vecsz_amd import make_params

    tickets = []
    for r in reqs:
        cap = int(r["x"].numel() * 12 + 4096)
        r["out_c"] = torch.empty(cap, dtype=torch.uint8, device="cuda")
        r["out_d"] = torch.empty_like(r["x"])
        torch.cuda.synchronize()
        tickets.append(("c", r, server.submit_compress(r["p"], r["x"].data_ptr(), r["out_c"].data_ptr(), cap)))
        pd = make_params((1,), 1e-4, "abs")
        tickets.append(("d", r, server.submit_decompress(pd, r["a"].data_ptr(), r["a"].numel(), r["out_d"].data_ptr(),
                                                         r["out_d"].numel())))
    for kind, r, t in tickets:
        res = server.wait(t)
        if kind == "c":
            assert res["out_bytes"] == r["a"].numel() and res["n_outlier"] == r["nout"]
            assert torch.equal(r["out_c"][: res["out_bytes"]], r["a"])
        else:
            assert torch.equal(r["out_d"].view(torch.int32), r["y"].view(torch.int32))


def test_server_results_match_standalone(gpu_ctx):
    from vecsz_amd import Server

    reqs = _requests(gpu_ctx)
    srv = Server(0, slots=4)
    print("mixed pass 1", flush=True)
    _run_mixed(srv, reqs)
    print("mixed pass 2", flush=True)
    _run_mixed(srv, reqs)  # resident slot scratch reused
    srv.close()


def test_server_budget_admission(gpu_ctx):
    from vecsz_amd import Server, VecszError, make_params

    reqs = _requests(gpu_ctx, seed=3)[:4]
    # room for about one request at a time: everything still completes, correctly
    srv = Server(0, slots=4, hbm_budget=64 << 20)
    print("tight budget", flush=True)
    _run_mixed(srv, reqs)
    # a request that can never fit is refused at submit
    x = torch.zeros(1 << 24, device="cuda")
    out = torch.empty(1 << 20, dtype=torch.uint8, device="cuda")
    with pytest.raises(VecszError):
        srv.submit_compress(make_params((1 << 24,), 1e-3, "abs"), x.data_ptr(), out.data_ptr(), out.numel())
    srv.close()


def test_server_folds_queued_compress_requests(gpu_ctx):
    """Small compress requests waiting together are folded into one fused
    batch launch (continuous batching); each archive is the one the request
    gives alone."""
    from vecsz_amd import Server, datagen as dg, make_params

    rng = np.random.default_rng(9)
    reqs = []
    for i in range(40):
        nd = int(rng.integers(1, 4))
        s = (int(rng.integers(100, 50000)),) if nd == 1 else tuple(int(v) for v in rng.integers(8, 300 if nd == 2 else 48, nd))
        x = torch.from_numpy(dg.grf(s, seed=50 + i) if nd > 1 else
                             np.cumsum(rng.normal(0, 0.05, s)).astype(np.float32)).cuda()
        eb = float(rng.choice([1e-2, 1e-3, 1e-4]))
        a, nout = gpu_ctx.compress(x, eb, "abs")
        cap = int(x.numel() * 12 + 4096)
        reqs.append({"x": x, "p": make_params(tuple(reversed(s)), eb, "abs"), "a": a.clone(), "nout": nout,
                     "out": torch.empty(cap, dtype=torch.uint8, device="cuda"), "cap": cap})
    torch.cuda.synchronize()
    srv = Server(0, slots=1)
    print("folded compress requests", flush=True)
    tickets = [srv.submit_compress(r["p"], r["x"].data_ptr(), r["out"].data_ptr(), r["cap"]) for r in reqs]
    res = [srv.wait(t) for t in tickets]
    srv.close()
    for r, q in zip(reqs, res):
        assert q["out_bytes"] == r["a"].numel() and q["n_outlier"] == r["nout"]
        assert torch.equal(r["out"][: q["out_bytes"]], r["a"])
    # requests that shared a launch started and ended together
    assert len(set(q["run_ms"] for q in res)) < len(res), [q["run_ms"] for q in res]


def test_server_folds_queued_decompress_requests(gpu_ctx):
    """Small decompress requests waiting together are folded into one fused
    batch decompress; each output is the one the request gives alone, and a
    member whose archive is corrupt fails on its own while the others
    succeed."""
    from vecsz_amd import Server, VecszError, datagen as dg, make_params

    rng = np.random.default_rng(21)
    reqs = []
    for i in range(40):
        nd = int(rng.integers(1, 4))
        s = (int(rng.integers(100, 50000)),) if nd == 1 else tuple(int(v) for v in rng.integers(8, 300 if nd == 2 else 48, nd))
        x = torch.from_numpy(dg.grf(s, seed=90 + i) if nd > 1 else
                             np.cumsum(rng.normal(0, 0.05, s)).astype(np.float32)).cuda()
        eb = float(rng.choice([1e-2, 1e-3, 1e-4]))
        a, _ = gpu_ctx.compress(x, eb, "abs")
        a = a.clone()
        y = gpu_ctx.decompress(a, "abs")
        reqs.append({"a": a, "y": y.clone(), "out": torch.empty_like(x)})
    bad = 17
    reqs[bad]["a"][0] = 9  # ndim 9: a corrupt header
    torch.cuda.synchronize()
    srv = Server(0, slots=1)
    print("folded decompress requests", flush=True)
    pd = make_params((1,), 1e-4, "abs")
    tickets = [srv.submit_decompress(pd, r["a"].data_ptr(), r["a"].numel(), r["out"].data_ptr(), r["out"].numel())
               for r in reqs]
    res = []
    for k, t in enumerate(tickets):
        if k == bad:
            with pytest.raises(VecszError):
                srv.wait(t)
            res.append(None)
        else:
            res.append(srv.wait(t))
    srv.close()
    for k, (r, q) in enumerate(zip(reqs, res)):
        if k != bad:
            assert torch.equal(r["out"].view(torch.int32), r["y"].view(torch.int32)), k
    # requests that shared a launch started and ended together
    runs = [q["run_ms"] for q in res if q is not None]
    assert len(set(runs)) < len(runs), runs


def test_server_fold_members_fail_alone(gpu_ctx):
    """A folded compress batch with one output buffer too small, and a folded
    decompress batch holding an archive with a pad section (which sends that
    fold to single calls): the short request fails alone, every other
    result equals the request's standalone result."""
    from vecsz_amd import Server, VecszError, datagen as dg, make_params

    xs = [torch.from_numpy(dg.grf((40, 48 + i, 56), seed=200 + i)).cuda() for i in range(12)]
    alone = [gpu_ctx.compress(x, 1e-3, "abs") for x in xs]
    outs = [torch.empty(int(x.numel() * 12 + 4096), dtype=torch.uint8, device="cuda") for x in xs]
    caps = [o.numel() for o in outs]
    short = 7
    caps[short] = alone[short][0].numel() - 1
    torch.cuda.synchronize()
    srv = Server(0, slots=1)
    print("folded compress with a short buffer", flush=True)
    tickets = [srv.submit_compress(make_params(tuple(reversed(x.shape)), 1e-3, "abs"), x.data_ptr(), o.data_ptr(), c)
               for x, o, c in zip(xs, outs, caps)]
    for k, t in enumerate(tickets):
        if k == short:
            with pytest.raises(VecszError):
                srv.wait(t)
        else:
            q = srv.wait(t)
            assert q["out_bytes"] == alone[k][0].numel() and torch.equal(outs[k][: q["out_bytes"]], alone[k][0]), k
    # decompress: archive 5 has a pad section
    archives = [a.clone() for a, _ in alone]
    archives[5] = gpu_ctx.compress(xs[5], 1e-3, "abs", pad_mode="block", pad_type="avg")[0].clone()
    want = [gpu_ctx.decompress(a, "abs") for a in archives]
    ys = [torch.empty_like(x) for x in xs]
    torch.cuda.synchronize()
    print("folded decompress with a padded archive", flush=True)
    pd = make_params((1,), 1e-4, "abs")
    tickets = [srv.submit_decompress(pd, a.data_ptr(), a.numel(), y.data_ptr(), y.numel()) for a, y in zip(archives, ys)]
    for t in tickets:
        srv.wait(t)
    srv.close()
    for k in range(len(xs)):
        assert torch.equal(ys[k].view(torch.int32), want[k].view(torch.int32)), k
