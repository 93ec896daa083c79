"""The served form on a Poisson request stream: 24 requests, half compress and
half decompress, sizes log-uniform from 64 KiB to 512 MiB of f32 (the largest
set to 512 MiB), 1D/2D/3D fields, abs eb 1e-4, submitted at exponential gaps
to a server with 8 slots and a 200 GB budget, so that requests arrive while
others run and overlap on the GPU.  Every result must equal what the same
request gives alone (vecsz_compress / vecsz_decompress on one context), and
the archives of requests up to 64 MiB must equal the oracle's."""
import concurrent.futures as cf
import time

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _shape(n, nd):
    if nd == 1:
        return (n,)
    if nd == 2:
        y = max(8, int(round(np.sqrt(n))))
        return (y, max(8, n // y))
    z = max(8, int(round(np.cbrt(n))))
    y = max(8, int(round(np.sqrt(n / z))))
    return (z, y, max(8, n // (z * y)))


def test_poisson_stream_results_equal_standalone(gpu_ctx):
    from vecsz_amd import Server, datagen, make_params

    rng = np.random.default_rng(21)
    sizes = np.exp(rng.uniform(np.log(64 << 10), np.log(512 << 20), size=24))
    sizes[int(np.argmax(sizes))] = 512 << 20
    reqs = []
    for i, nbytes in enumerate(sizes):
        nd = int(rng.integers(1, 4))
        shape = _shape(max(1, int(nbytes // 4)), nd)
        x = datagen.grf(shape, seed=400 + i, device="cuda", pow2=True) if nd > 1 else \
            datagen.walk_spikes_1d(shape[0], seed=400 + i, device="cuda")
        p = make_params(tuple(reversed(shape)), 1e-4, "abs")
        a, nout = gpu_ctx.compress(x, 1e-4, "abs")  # alone
        r = {"kind": "c" if i % 2 == 0 else "d", "x": x, "p": p, "a": a.clone(), "nout": nout}
        if r["kind"] == "c":
            r["cap"] = gpu_ctx.compress_bound(p)
            r["out"] = torch.empty(r["cap"], dtype=torch.uint8, device="cuda")
        else:
            r["y"] = gpu_ctx.decompress(a, "abs").clone()  # alone
            r["out"] = torch.empty_like(x)
        reqs.append(r)
    torch.cuda.synchronize()
    total = sum(4 * r["x"].numel() for r in reqs)
    print(f"stream: {len(reqs)} requests, {total / 2**30:.2f} GiB, largest {max(sizes) / 2**20:.0f} MiB", flush=True)

    # the standalone archives of the smaller requests against the oracle
    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        small = [r for r in reqs if 4 * r["x"].numel() <= (64 << 20)]
        futs = [(r, ex.submit(po.compress, r["x"].cpu().numpy(), 1e-4, False)) for r in small]

        pd = make_params((1,), 1e-4, "abs")
        srv = Server(0, slots=8, hbm_budget=200 * 10**9)
        gaps = np.random.default_rng(3).exponential(2e-3, size=len(reqs))
        order = np.random.default_rng(4).permutation(len(reqs))
        tickets = []
        t0 = time.perf_counter()
        for k, i in enumerate(order):
            time.sleep(float(gaps[k]))
            r = reqs[i]
            if r["kind"] == "c":
                t = srv.submit_compress(r["p"], r["x"].data_ptr(), r["out"].data_ptr(), r["cap"])
            else:
                t = srv.submit_decompress(pd, r["a"].data_ptr(), r["a"].numel(), r["out"].data_ptr(), r["out"].numel())
            tickets.append((r, t, (time.perf_counter() - t0) * 1e3))
        res = [(r, srv.wait(t), ts) for r, t, ts in tickets]
        span = time.perf_counter() - t0
        srv.close()
        # run intervals [submit + queue, submit + latency]: some requests ran
        # at the same time as others
        iv = sorted((ts + q["queue_ms"], ts + q["latency_ms"]) for _, q, ts in res)
        overlapped = sum(1 for a, b in zip(iv, iv[1:]) if b[0] < a[1])
        print(f"stream served in {span * 1e3:.1f} ms; {overlapped} overlapping starts", flush=True)
        res = [(r, q) for r, q, _ in res]
        for r, q in res:  # (a request that failed raised in wait)
            if r["kind"] == "c":
                assert q["out_bytes"] == r["a"].numel() and q["n_outlier"] == r["nout"]
                assert torch.equal(r["out"][: q["out_bytes"]], r["a"])
            else:
                assert torch.equal(r["out"].view(torch.int32), r["y"].view(torch.int32))
        assert overlapped > 0
        for r, f in futs:
            want, nout, _ = f.result()
            assert r["a"].cpu().numpy().tobytes() == want and r["nout"] == nout
