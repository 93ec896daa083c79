#!/usr/bin/env python3
"""Throughput of the dual-quantization + Huffman path on one MI355X.

Workload (default): 3D float32 field 512 x 512 x 1024 (x first; 1 GiB),
Gaussian random field with E(k) ~ k^-5/3 synthesised on the GPU from a seed,
scaled to [-50, 50], compressed at r2r eb = 1e-4 with 8^3 Lorenzo blocks and a
4096-code dictionary.  A step is one compress of the field (input resident in
HBM, archive written to HBM); decompression of the resulting archive is timed
the same way and reported beside it.  The 512^3 field at the same settings is
measured after it and reported under "grf512_r2r1e-4".

Prints ONE JSON line on stdout; progress goes to stderr.
    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GB/s compress/decompress (1 MI355X, eb=1e-4) vs CPU; %HBM peak; p99 lat"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_avg(ctx, fn, reps):
    """Per-call device time of each stage (a stage that runs several times in
    one call, like the decoder's re-decode passes, is summed), averaged."""
    acc = {}
    for _ in range(reps):
        fn()
        per = {}
        for name, ms in ctx.timings():
            per[name] = per.get(name, 0.0) + ms
        for name, ms in per.items():
            acc.setdefault(name, []).append(ms)
    return {k: sum(v) / len(v) for k, v in acc.items()}


def dq_kernel_name(ndim, block):
    return {3: f"k_dualquant_3d_x4<{block}, true>", 2: f"k_dualquant_2d_x4<{block}, true>",
            1: f"k_dualquant_1d<{block}, false>"}[ndim]


def measured_traffic(kernel, dims, eb):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE and WRITE_SIZE in KiB per dispatch,
    scaled by what the same counters read for a plain copy of known size);
    None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_" + "x".join(map(str, dims)) + f"_r2r{eb:g}.md")
    if not os.path.exists(path):
        return None
    fetch_scale = write_scale = None
    rows = {}
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 2 and cells[0] == "fetch_scale":
            fetch_scale = float(cells[1])
        elif len(cells) >= 2 and cells[0] == "write_scale":
            write_scale = float(cells[1])
        elif len(cells) >= 3 and cells[0].startswith("k_"):
            try:
                rows[cells[0]] = (float(cells[1]), float(cells[2]))
            except ValueError:
                pass
    if kernel not in rows:
        # a kernel whose template argument the library picks at run time
        # (k_pack<book>): the one instantiation the summary holds
        alike = [k for k in rows if k.split("<")[0] == kernel.split("<")[0]]
        kernel = alike[0] if len(alike) == 1 else kernel
    if kernel not in rows or fetch_scale is None or write_scale is None:
        return None
    f, w = rows[kernel]
    return int(round((f * fetch_scale + w * write_scale) * 1024))


def host_threads():
    """Host threads this process may use: OMP_NUM_THREADS when set (the GPU
    box sets it to its CPU share), else the CPUs in the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, len(os.sched_getaffinity(0)))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_cpu(po, x_host, eb_abs, dict_size, block, threads, ref):
    if ref:
        po.ref().ref_set_threads(threads)
        t0 = time.perf_counter()
        a, _ = po.ref_compress(x_host, eb_abs, False, dict_size, block, threads=threads)
        tc = time.perf_counter() - t0
        t0 = time.perf_counter()
        po.ref_decompress(a, None, False, dict_size, block)
        td = time.perf_counter() - t0
    else:
        t0 = time.perf_counter()
        a, _, _ = po.compress(x_host, eb_abs, False, dict_size, block)
        tc = time.perf_counter() - t0
        t0 = time.perf_counter()
        po.decompress(a, x_host.shape, False, 0.0, dict_size, block)
        td = time.perf_counter() - t0
    return tc, td


def cpu_baseline(x_host, eb_abs, dims, dict_size, block):
    """The reference's own Compress/Decompress (oracle/_ref, built from its
    sources in the CPU container and shipped with the tree) when present,
    else the C restatement (one thread), timed on this host: the whole field
    on all of this process's threads, and a slab of it (the first quarter of
    its slowest axis) on one thread."""
    import numpy as np

    from oracle import pyoracle as po

    ref = po.ref_available()
    threads = host_threads() if ref else 1
    # abs mode with eb_final: the same archive bits as r2r without the
    # reference's re-read of the input file
    tc, td = _time_cpu(po, x_host, eb_abs, dict_size, block, threads, ref)
    slab = np.ascontiguousarray(x_host[: max(1, x_host.shape[0] // 4)])
    tc1, td1 = _time_cpu(po, slab, eb_abs, dict_size, block, 1, ref)
    n, n1 = x_host.size, slab.size
    return {
        "value": round(4 * n / tc / 1e9, 4),
        "unit": "GB/s",
        "cores": threads,
        "kind": "reference" if ref else "port",
        "sample": f"one compress and one decompress of the full {'x'.join(map(str, dims))} field in abs mode at the "
                  f"same eb_final (no file IO) on {threads} OpenMP threads; the 1-thread figures are one compress and "
                  f"one decompress of its first {slab.shape[0]} slowest-axis planes ({4 * n1 / 2**20:.0f} MiB)",
        "cpu_model": cpu_model(),
        "compress_s": round(tc, 3),
        "decompress_s": round(td, 3),
        "decompress_gbps": round(4 * n / td / 1e9, 4),
        "one_thread": {"compress_gbps": round(4 * n1 / tc1 / 1e9, 4), "decompress_gbps": round(4 * n1 / td1 / 1e9, 4),
                       "compress_s": round(tc1, 3), "decompress_s": round(td1, 3)},
    }


def apply_launch_shape(ctx, args, x, eb, mode):
    """Launch shapes for the run: given (--launch-shape), none (--no-autotune)
    or tuned on the field x by vecsz_autotune (archives are the same either
    way).  Returns what was applied, for the JSON line."""
    if args.launch_shape:
        dq, enc = (int(v) for v in args.launch_shape.split(","))
        ctx.set_launch_shape(dq, enc)
        return {"dq_grid": dq, "enc_grid": enc, "source": "given"}
    if args.no_autotune:
        return {"dq_grid": 768, "enc_grid": 2048, "source": "default"}
    log("autotuning launch shapes")
    t = ctx.autotune(x, eb, mode, args.block, args.dict)
    log("autotune:", t)
    return {"dq_grid": t["dq_grid"], "enc_grid": t["enc_grid"], "source": "autotune",
            "dq_ms": round(t["dq_ms"], 4), "enc_ms": round(t["enc_ms"], 4), "fastest_block": t["block"],
            "ms_by_block": {str(k): round(v, 4) for k, v in t["ms_by_block"].items()},
            "bytes_by_block": {str(k): v for k, v in t["bytes_by_block"].items()}}


def bench_batch(args):
    """The ~8 GiB mix of 96 fields of uneven shape (2D and 3D GRFs, per-field
    r2r eb; datagen.batch_mix): one batched call against the same fields
    compressed by one call each, on one context; archives must be identical."""
    import torch

    from vecsz_amd import Context, datagen, make_params

    torch.cuda.set_device(0)
    mix, mix_seed = datagen.batch_mix(96, seed=0)  # ~8 GiB
    fields, params, outs = [], [], []
    for k, (shape, eb) in enumerate(mix):
        x = datagen.grf(shape, seed=100 + k, device="cuda", pow2=True)
        fields.append(x)
        if k % 16 == 15:
            torch.cuda.synchronize()
            log(f"generated {k + 1} fields")
        p = make_params(tuple(reversed(shape)), eb, "r2r", args.block, args.dict)
        params.append(p)
    ctx = Context(0)
    big = max(range(len(fields)), key=lambda k: fields[k].numel())
    tune = apply_launch_shape(ctx, args, fields[big], mix[big][1], "r2r")
    caps = [ctx.compress_bound(p) for p in params]
    outs = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    outs1 = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    torch.cuda.synchronize()
    total = sum(4 * f.numel() for f in fields)
    log(f"c4 mix: {len(fields)} fields, {total / 2**30:.3f} GiB")

    def plain():
        return [ctx.compress_into(p, f.data_ptr(), o.data_ptr(), o.numel()) for p, f, o in zip(params, fields, outs1)]

    def batch():
        return ctx.compress_batch_into(params, [f.data_ptr() for f in fields], [o.data_ptr() for o in outs], caps)

    for _ in range(args.warmup):
        log("warmup: per-field calls")
        plain()
        log("warmup: batched call")
        batch()
    torch.cuda.synchronize()
    log("timing compress")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r1 = plain()
    torch.cuda.synchronize()
    tp = (time.perf_counter() - t0) / args.steps
    sweep = {}
    for nw in reversed([int(w) for w in args.workers.split(",")]):
        ctx.set_workers(nw)
        batch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rb = batch()
        torch.cuda.synchronize()
        sweep[nw] = (time.perf_counter() - t0) / args.steps
        log(f"batch with {nw} workers: {sweep[nw] * 1e3:.3f} ms")
    tb = sweep[int(args.workers.split(",")[0])]
    same = r1 == rb and all(torch.equal(a[:s], b[:s]) for a, b, (s, _) in zip(outs, outs1, rb))
    # the worker-context path on the same batch (the fused path is the default)
    ctx.set_workers(int(args.workers.split(",")[0]))
    ctx.set_batch_fused(False)
    batch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rw = batch()
    torch.cuda.synchronize()
    tw = (time.perf_counter() - t0) / args.steps
    ctx.set_batch_fused(True)
    log(f"worker-context batch: {tw * 1e3:.3f} ms")
    same = same and rw == rb and all(torch.equal(a[:s], b[:s]) for a, b, (s, _) in zip(outs, outs1, rb))

    # decompress both ways
    vr = [float((f.max() - f.min()).item()) for f in fields]
    dparams = [make_params((1,), 1e-4, "r2r", args.block, args.dict, v) for v in vr]
    ys = [torch.empty_like(f) for f in fields]
    ys1 = [torch.empty_like(f) for f in fields]

    def dplain():
        for p, a, (s, _), y in zip(dparams, outs, rb, ys1):
            ctx.decompress_into(p, a.data_ptr(), s, y.data_ptr(), y.numel())

    def dbatch():
        ctx.decompress_batch_into(dparams, [a.data_ptr() for a in outs], [s for s, _ in rb],
                                  [y.data_ptr() for y in ys], [y.numel() for y in ys])

    log("timing decompress")
    for _ in range(args.warmup):
        dplain()
        dbatch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dplain()
    torch.cuda.synchronize()
    tdp = (time.perf_counter() - t0) / args.steps
    # the batch call as it runs by default (these fields average ~85 MB: the
    # worker contexts), over set_workers values
    dsweep = {}
    for nw in reversed([int(w) for w in args.workers.split(",")] + [1, 2, 4]):
        if nw in dsweep:
            continue
        ctx.set_workers(nw)
        dbatch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dbatch()
        torch.cuda.synchronize()
        dsweep[nw] = (time.perf_counter() - t0) / args.steps
        log(f"decompress batch with {nw} workers: {dsweep[nw] * 1e3:.3f} ms")
    ctx.set_workers(int(args.workers.split(",")[0]))
    tdb = dsweep[int(args.workers.split(",")[0])]
    dsame = all(torch.equal(a.view(torch.int32), b.view(torch.int32)) for a, b in zip(ys, ys1))
    # the fused path forced on the same archives
    log("timing the fused decompress batch")
    ctx.set_batch_fused("always")
    ys.clear()
    ys.extend(torch.empty_like(f) for f in fields)
    dbatch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dbatch()
    torch.cuda.synchronize()
    tdf = (time.perf_counter() - t0) / args.steps
    dsame = dsame and all(torch.equal(a.view(torch.int32), b.view(torch.int32)) for a, b in zip(ys, ys1))
    # stage times of the fused batch calls (HIP events on the context's stream)
    ctx.enable_timing(True)
    dbst = stage_avg(ctx, dbatch, min(args.steps, 3))
    ctx.set_batch_fused(True)
    bst = stage_avg(ctx, batch, min(args.steps, 3))
    ctx.enable_timing(False)
    log("fused batch compress stages (ms):", {k: round(v, 4) for k, v in bst.items()})
    log("fused batch decompress stages (ms):", {k: round(v, 4) for k, v in dbst.items()})
    arch = sum(s for s, _ in rb)
    del ys, ys1
    torch.cuda.empty_cache()
    small = bench_small_batch(args, ctx)
    out = {
        "metric": METRIC + " [batched form, 96-field mix]",
        "value": round(total / tb / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tb * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (96 GRFs E(k)~k^-5/3 cropped from power-of-two grids, seeds 100..195, shapes and error "
                f"bounds from datagen.batch_mix(96, 0), draw seed {mix_seed})",
        "config": {"workload": f"batched mix: {len(fields)} fields, {total / 2**30:.3f} GiB, r2r eb per field in "
                               "{1e-2,1e-3,1e-4}, B 8, dict 4096", "parallelism": "single-gpu"},
        "per_field_calls_ms": round(tp * 1e3, 3),
        "batch_ms_by_workers": {str(k): round(v * 1e3, 3) for k, v in sorted(sweep.items())},
        "batch_speedup": round(tp / tb, 3),
        "batch_worker_contexts_ms": round(tw * 1e3, 3),
        "launch_shape": tune,
        "identical_archives": bool(same),
        "decompress_gbps": round(total / tdb / 1e9, 2),
        "decompress_per_field_calls_ms": round(tdp * 1e3, 3),
        "decompress_batch_ms": round(tdb * 1e3, 3),
        "decompress_batch_speedup": round(tdp / tdb, 3),
        "decompress_batch_ms_by_workers": {str(k): round(v * 1e3, 3) for k, v in sorted(dsweep.items())},
        "decompress_batch_fused_ms": round(tdf * 1e3, 3),
        "batch_stages_ms": {"compress": {k: round(v, 4) for k, v in bst.items()},
                            "decompress": {k: round(v, 4) for k, v in dbst.items()}},
        "identical_outputs": bool(dsame),
        "archive_bytes": arch,
        "small_fields": small,
    }
    print(json.dumps(out), flush=True)


def bench_small_batch(args, ctx, nfields=256, seed=1):
    """A batch of small fields beside the C4 mix: 256 GRFs of 16K..1M
    elements (log-uniform; half 2D, half 3D; r2r eb per field from {1e-2,
    1e-3, 1e-4}).  Per-field calls against the batch calls as they run by
    default (both fused for fields this small) and against the worker
    contexts; results must be identical."""
    import numpy as np
    import torch

    from vecsz_amd import datagen, make_params

    rng = np.random.default_rng(seed)
    fields, params, ebs = [], [], []
    for k in range(nfields):
        n = float(np.exp(rng.uniform(np.log(16 << 10), np.log(1 << 20))))
        if k % 2 == 0:
            y = max(8, int(round(np.sqrt(n))))
            shape = (y, max(8, int(round(n / y))))
        else:
            z = max(8, int(round(np.cbrt(n))))
            shape = (z, z, max(8, int(round(n / (z * z)))))
        eb = float(rng.choice([1e-2, 1e-3, 1e-4]))
        fields.append(datagen.grf(shape, seed=1000 + k, device="cuda", pow2=True))
        params.append(make_params(tuple(reversed(shape)), eb, "r2r", args.block, args.dict))
        ebs.append(eb)
    torch.cuda.synchronize()
    total = sum(4 * f.numel() for f in fields)
    log(f"small-field batch: {nfields} fields, {total / 2**20:.1f} MiB")
    caps = [ctx.compress_bound(p) for p in params]
    outs = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    outs1 = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps, r

    plain = lambda: [ctx.compress_into(p, f.data_ptr(), o.data_ptr(), o.numel())
                     for p, f, o in zip(params, fields, outs1)]
    batch = lambda: ctx.compress_batch_into(params, [f.data_ptr() for f in fields], [o.data_ptr() for o in outs], caps)
    tp, r1 = timed(plain)
    tb, rb = timed(batch)  # the batch replays as two captured graphs from the third call on
    ctx.set_batch_graph(False)
    tbe, rbe = timed(batch)
    ctx.set_batch_graph(True)
    ctx.set_batch_fused(False)
    tw, rw = timed(batch)
    ctx.set_batch_fused(True)
    same = r1 == rb == rw == rbe and all(torch.equal(a[:s], b[:s]) for a, b, (s, _) in zip(outs, outs1, rb))
    vr = [float((f.max() - f.min()).item()) for f in fields]
    dparams = [make_params((1,), 1e-4, "r2r", args.block, args.dict, v) for v in vr]
    ys = [torch.empty_like(f) for f in fields]
    ys1 = [torch.empty_like(f) for f in fields]
    dplain = lambda: [ctx.decompress_into(p, a.data_ptr(), s, y.data_ptr(), y.numel())
                      for p, a, (s, _), y in zip(dparams, outs, rb, ys1)]
    dbatch = lambda: ctx.decompress_batch_into(dparams, [a.data_ptr() for a in outs], [s for s, _ in rb],
                                               [y.data_ptr() for y in ys], [y.numel() for y in ys])
    tdp, _ = timed(dplain)
    tdb, _ = timed(dbatch)
    dsame = all(torch.equal(a.view(torch.int32), b.view(torch.int32)) for a, b in zip(ys, ys1))
    ctx.set_batch_fused(False)
    tdw, _ = timed(dbatch)
    ctx.set_batch_fused(True)
    dsame = dsame and all(torch.equal(a.view(torch.int32), b.view(torch.int32)) for a, b in zip(ys, ys1))
    res = {"fields": nfields, "bytes": total,
           "compress_ms": {"per_field_calls": round(tp * 1e3, 3), "batch_fused": round(tb * 1e3, 3),
                           "batch_fused_no_graph": round(tbe * 1e3, 3),
                           "batch_worker_contexts": round(tw * 1e3, 3)},
           "decompress_ms": {"per_field_calls": round(tdp * 1e3, 3), "batch_fused": round(tdb * 1e3, 3),
                             "batch_worker_contexts": round(tdw * 1e3, 3)},
           "identical_archives": bool(same), "identical_outputs": bool(dsame)}
    log("small-field batch:", res)
    return res


def bench_serve(args):
    """Served form: a Poisson stream of compress and decompress requests (50/50,
    sizes log-uniform 64 KiB..512 MiB, 1D/2D/3D, abs eb 1e-4) offered at a
    fraction of the measured peak to the request server (200 GB budget);
    p50/p99 latency and sustained GB/s.  Peak = the same requests submitted
    all at once."""
    import random

    import numpy as np
    import torch

    from vecsz_amd import Context, Server, datagen, make_params

    torch.cuda.set_device(0)
    rng = np.random.default_rng(5)
    ctx = Context(0)
    reqs = []
    for i in range(args.requests):
        nbytes = float(np.exp(rng.uniform(np.log(64 << 10), np.log(512 << 20))))
        n = max(1, int(nbytes // 4))
        nd = int(rng.integers(1, 4))
        if nd == 1:
            shape = (n,)
        elif nd == 2:
            y = max(8, int(round(np.sqrt(n))))
            shape = (y, max(8, n // y))
        else:
            z = max(8, int(round(np.cbrt(n))))
            y = max(8, int(round(np.sqrt(n / z))))
            shape = (z, y, max(8, n // (z * y)))
        x = datagen.grf(shape, seed=300 + i, device="cuda", pow2=True) if nd > 1 else \
            datagen.walk_spikes_1d(shape[0], seed=300 + i, device="cuda")
        kind = "c" if rng.random() < 0.5 else "d"
        p = make_params(tuple(reversed(shape)), 1e-4, "abs")
        cap = ctx.compress_bound(p)
        r = {"kind": kind, "x": x, "p": p, "bytes": 4 * x.numel()}
        if kind == "c":
            r["out"] = torch.empty(cap, dtype=torch.uint8, device="cuda")
        else:
            a = torch.empty(cap, dtype=torch.uint8, device="cuda")
            size, _ = ctx.compress_into(p, x.data_ptr(), a.data_ptr(), cap)
            r["a"], r["size"] = a, size
            r["out"] = torch.empty_like(x)
        reqs.append(r)
        if i % 16 == 15:
            torch.cuda.synchronize()
            log(f"prepared {i + 1} requests")
    torch.cuda.synchronize()
    total = sum(r["bytes"] for r in reqs)
    pd = make_params((1,), 1e-4, "abs")
    srv = Server(0, slots=8, hbm_budget=200 * 10**9)

    def submit(r):
        if r["kind"] == "c":
            return srv.submit_compress(r["p"], r["x"].data_ptr(), r["out"].data_ptr(), r["out"].numel())
        return srv.submit_decompress(pd, r["a"].data_ptr(), r["size"], r["out"].data_ptr(), r["out"].numel())

    # warm the slots, then the peak: everything submitted at once
    for t in [submit(r) for r in reqs]:
        srv.wait(t)
    t0 = time.perf_counter()
    for t in [submit(r) for r in reqs]:
        srv.wait(t)
    peak = total / (time.perf_counter() - t0)
    log(f"peak (closed loop): {peak / 1e9:.2f} GB/s over {total / 2**30:.2f} GiB")

    # open loop: Poisson arrivals at load * peak (bytes/s); submission order shuffled
    lam = args.load * peak / (total / len(reqs))
    order = list(range(len(reqs)))
    random.Random(7).shuffle(order)
    gaps = np.random.default_rng(11).exponential(1.0 / lam, size=len(reqs))
    arrivals = np.cumsum(gaps)
    tickets = []
    start = time.perf_counter()
    for k, i in enumerate(order):
        delay = arrivals[k] - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        tickets.append(submit(reqs[i]))
    res = [srv.wait(t) for t in tickets]
    span = time.perf_counter() - start
    lat = np.array([r["latency_ms"] for r in res])
    qms = np.array([r["queue_ms"] for r in res])
    srv.close()
    out = {
        "metric": METRIC + " [served form, Poisson 50/50 compress/decompress]",
        "value": round(total / span / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": 1,
        "ms_per_step": round(span * 1e3 / len(reqs), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (GRFs cropped from power-of-two grids for 2D/3D, random walk + spikes for 1D)",
        "config": {"workload": f"served: {len(reqs)} requests, {total / 2**30:.2f} GiB, sizes log-uniform 64 KiB..512 MiB, "
                               f"abs eb 1e-4, offered load {args.load} x peak, 8 slots, 200 GB budget",
                   "parallelism": "single-gpu"},
        "peak_gbps": round(peak / 1e9, 2),
        "offered_gbps": round(args.load * peak / 1e9, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3)},
        "queue_ms": {"p50": round(float(np.percentile(qms, 50)), 3), "p99": round(float(np.percentile(qms, 99)), 3)},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shape", default="512,512,512", help="field workload: C order (z,y,x)")
    ap.add_argument("--eb", type=float, default=1e-4)
    ap.add_argument("--block", type=int, default=8)
    ap.add_argument("--dict", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="target", choices=["target", "grf512", "field", "c1", "c3", "c4", "c5"],
                    help="target (default line): 3D GRF 512x512x1024 (1 GiB), r2r 1e-4, with the 512^3 "
                         "configuration reported beside it; grf512: 3D GRF 512^3, r2r 1e-4; field: 3D GRF of "
                         "--shape at r2r --eb; c1: 2D 1800x3600 sinusoids, abs 1e-4; c3: 1D 2^28 walk + spikes, "
                         "r2r 1e-4; c4: the ~8 GiB 96-field batched mix vs per-field calls; c5: served Poisson "
                         "request stream")
    ap.add_argument("--requests", type=int, default=96, help="c5: requests in the stream")
    ap.add_argument("--load", type=float, default=0.7, help="c5: offered load as a fraction of measured peak")
    ap.add_argument("--workers", default="8", help="c4: comma-separated worker counts to time (first is reported)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the launch-shape autotuner (shapes from --launch-shape, else the defaults)")
    ap.add_argument("--launch-shape", default="", help="DQ_GRID,ENC_GRID: set launch shapes instead of tuning")
    ap.add_argument("--no-extra", action="store_true",
                    help="target: skip the 512^3 configuration measured after it (profiling runs)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain device copy of the field (known bytes) to scale HBM counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a field would shard by slabs of "
              "Lorenzo blocks, with one histogram all-reduce). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    from vecsz_amd import Context, datagen, make_params

    if args.workload == "c4":
        return bench_batch(args)
    if args.workload == "c5":
        return bench_serve(args)

    torch.cuda.set_device(0)
    if args.workload == "target":
        # the 1 GiB target field is the line; the 512^3 configuration rides along
        out = measure_single(args, "target", full=True)
        if not args.no_extra:
            log("measuring the 512^3 configuration")
            torch.cuda.empty_cache()
            extra = measure_single(args, "grf512", full=False)
            out["grf512_r2r1e-4"] = {k: extra[k] for k in ("value", "ms_per_step", "decompress_gbps", "decompress_ms",
                                                              "archive_bytes", "pipeline_hbm_frac", "roofline",
                                                              "stages_ms", "latency_ms", "decode_index")}
    else:
        out = measure_single(args, args.workload, full=True)
    print(json.dumps(out), flush=True)


def make_field(args, workload):
    """(tensor, C-order shape, mode, eb, description) of a single-field workload."""
    from vecsz_amd import datagen
    import torch

    if workload == "c1":
        log("generating the 2D sinusoid field on the host")
        x = torch.from_numpy(datagen.sinusoids_2d(1800, 3600, seed=0)).cuda()
        return x, (1800, 3600), "abs", 1e-4, "2D f32 3600x1800 sinusoids + noise, abs eb=1e-4"
    if workload == "c3":
        log("generating the 1D walk + spikes on the GPU")
        x = datagen.walk_spikes_1d(1 << 28, seed=0, device="cuda")
        return x, (1 << 28,), "r2r", 1e-4, "1D f32 2^28 random walk + 0.1% spikes, r2r eb=1e-4"
    if workload == "target":
        shape, eb = (1024, 512, 512), 1e-4
    elif workload == "grf512":
        shape, eb = (512, 512, 512), 1e-4
    else:  # "field": --shape / --eb
        shape, eb = tuple(int(v) for v in args.shape.split(",")), args.eb
    log(f"generating GRF {shape} on the GPU")
    x = datagen.grf(shape, seed=0, device="cuda")
    what = f"{len(shape)}D f32 {'x'.join(map(str, reversed(shape)))} GRF r2r eb={eb}"
    return x, shape, "r2r", eb, what


def measure_single(args, workload, full=True):
    """One single-field workload: compress and decompress steps, per-stage
    times, the dominant kernel's roofline, and (full) the decode-index path,
    the one-pass encode for comparison and the CPU baseline."""
    import torch

    from vecsz_amd import Context, make_params

    x, shape, mode, eb, what = make_field(args, workload)
    dims = tuple(reversed(shape))
    torch.cuda.synchronize()
    n = x.numel()
    value_range = float((x.max() - x.min()).item()) if mode == "r2r" else 0.0

    ctx = Context(0)
    tune = apply_launch_shape(ctx, args, x, eb, mode)
    p = make_params(dims, eb, mode, args.block, args.dict)
    cap = ctx.compress_bound(p)
    arch = torch.empty(cap, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    pd = make_params((1,), 1e-4, mode, args.block, args.dict, value_range)

    def comp():
        return ctx.compress_into(p, x.data_ptr(), arch.data_ptr(), cap)

    size, nout = comp()

    def decomp():
        ctx.decompress_into(pd, arch.data_ptr(), size, y.data_ptr(), n)

    if args.calib_copy:
        y.copy_(x)  # reads and writes exactly 4*N bytes: the yardstick for FETCH_SIZE / WRITE_SIZE
        torch.cuda.synchronize()

    # warmup
    for _ in range(args.warmup):
        comp()
        decomp()
    torch.cuda.synchronize()

    # timed compress steps
    log(f"timing {args.steps} compress steps: {what}")
    ctx.enable_timing(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        size, nout = comp()
    torch.cuda.synchronize()
    tc = (time.perf_counter() - t0) / args.steps
    lat = []
    for _ in range(min(args.steps, 20)):
        s = time.perf_counter()
        comp()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - s)
    lat.sort()

    # timed decompress steps
    log(f"timing {args.steps} decompress steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        decomp()
    torch.cuda.synchronize()
    td = (time.perf_counter() - t0) / args.steps

    # the same with a decode index (a side product of compress kept beside
    # the archive): compress writing it, decompress using it
    iw = ctx.index_words(p)
    idx = torch.empty(iw, dtype=torch.int64, device="cuda")
    yi = torch.empty_like(x)

    def comp_idx():
        return ctx.compress_into(p, x.data_ptr(), arch.data_ptr(), cap, idx.data_ptr(), iw)

    def decomp_idx():
        ctx.decompress_into(pd, arch.data_ptr(), size, yi.data_ptr(), n, idx.data_ptr(), iw)

    for _ in range(args.warmup):
        comp_idx()
        decomp_idx()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        comp_idx()
    torch.cuda.synchronize()
    tci = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        decomp_idx()
    torch.cuda.synchronize()
    tdi = (time.perf_counter() - t0) / args.steps
    same_idx = bool(torch.equal(yi.view(torch.int32), y.view(torch.int32)))

    # correctness: the round trip must respect the error bound up to the
    # reference's fp32 slack (bit-exact parity is tested in tests/).
    err = (x.double() - y.double()).abs()
    slack = x.double().abs() * 2.0 ** -22
    eb_final = eb * value_range if mode == "r2r" else eb
    bound_ok = bool((err <= eb_final + slack).all().item())
    del err, slack

    # per-stage device time: HIP events on the library's stream, averaged over
    # a second run of the same K compress steps
    ctx.enable_timing(True)
    cst = stage_avg(ctx, comp, args.steps)
    cst1 = cstd = None
    tdev = 0.0
    if full:
        ctx.set_fused_encode(True)  # the one-pass encode, for comparison (same archive)
        cst1 = stage_avg(ctx, comp, args.steps)
        ctx.set_fused_encode(False)
        # the codebook built on the device (no host stop), for comparison (same archive)
        ctx.set_device_codebook(True)
        cstd = stage_avg(ctx, comp, min(args.steps, 5))
        ctx.enable_timing(False)
        comp()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(min(args.steps, 5)):
            comp()
        torch.cuda.synchronize()
        tdev = (time.perf_counter() - t0) / min(args.steps, 5)
        ctx.enable_timing(True)
        ctx.set_device_codebook(False)
    dst = stage_avg(ctx, decomp, args.steps)
    dsti = stage_avg(ctx, decomp_idx, args.steps)
    ctx.enable_timing(False)
    log("compress stages (ms):", {k: round(v, 4) for k, v in cst.items()})
    log("decompress stages (ms):", {k: round(v, 4) for k, v in dst.items()})

    # Dominant compress kernel and its algorithmic bytes per launch: the
    # range and dual-quantization kernels each read the field once (4 B per
    # element); the packer writes the archive (A bytes).  The u16 code array
    # they pass along is scratch, counted in `traffic`, not in `achieved`.
    kern = {"range": ("k_range", 4 * n), "dualquant": (dq_kernel_name(len(dims), args.block), 4 * n),
            "chunk_bits": ("k_chunk_bits<true>", 0), "scan": ("k_scan_local", 0), "pack": ("k_pack", size),
            "encode": ("k_encode", size)}
    dom = max((k for k in cst if k in kern), key=lambda k: cst[k])
    kname, abytes = kern[dom]
    achieved = abytes / (cst[dom] * 1e-3) / 1e9
    traffic = measured_traffic(kname, tuple(dims), eb)
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": kname,
        "alg_bytes": abytes,
        "avg_ms": round(cst[dom], 4),
    }

    comp_gbps = 4 * n / tc / 1e9
    dec_gbps = 4 * n / td / 1e9
    stages = {"compress": {k: round(v, 4) for k, v in cst.items()},
              "decompress": {k: round(v, 4) for k, v in dst.items()}}
    if cst1 is not None:
        stages["compress_one_pass_encode"] = {k: round(v, 4) for k, v in cst1.items()}
    if cstd is not None:
        stages["compress_device_codebook"] = {k: round(v, 4) for k, v in cstd.items()}
    out = {
        "metric": METRIC,
        "value": round(comp_gbps, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tc * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic, seed 0 (see vecsz_amd/datagen.py)",
        "config": {"workload": what + " compress",
                   "dims_xyz": list(dims), "eb": eb, "mode": mode, "block": args.block,
                   "dict_size": args.dict, "parallelism": "single-gpu"},
        "decompress_gbps": round(dec_gbps, 2),
        "decompress_ms": round(td * 1e3, 4),
        "decode_index": {"words": iw, "bytes": 8 * iw, "compress_gbps": round(4 * n / tci / 1e9, 2),
                         "decompress_gbps": round(4 * n / tdi / 1e9, 2), "decompress_ms": round(tdi * 1e3, 4),
                         "identical_output": same_idx,
                         "decompress_stages_ms": {k: round(v, 4) for k, v in dsti.items()}},
        "archive_bytes": size,
        "compression_ratio": round(4 * n / size, 4),
        "n_outlier": nout,
        "error_bound_ok": bound_ok,
        "latency_ms": {"p50": round(lat[len(lat) // 2] * 1e3, 4), "p99": round(lat[-1] * 1e3, 4)},
        "stages_ms": stages,
        "roofline": roofline,
        "launch_shape": tune,
        "pipeline_hbm_frac": {"compress": round((4 * n + size) / tc / 1e9 / HBM_PEAK_GBS, 4),
                              "decompress": round((4 * n + size) / td / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if full:
        out["codebook_place"] = {"host_compress_ms": round(tc * 1e3, 4), "device_compress_ms": round(tdev * 1e3, 4),
                                 "host_codebook_ms": round(cst.get("host_codebook", 0.0), 4),
                                 "device_codebook_ms": round(cstd.get("codebook", 0.0), 4)}
    if full and not args.no_cpu:
        log("timing the CPU baseline")
        x_host = x.cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(x_host, eb_final, dims, args.dict, args.block)
        out["speedup_vs_cpu"] = round(comp_gbps / out["cpu_baseline"]["value"], 2)
        del x_host
    elif full:
        out["cpu_baseline"] = None
    del x, y, yi, arch, idx, ctx
    return out


if __name__ == "__main__":
    main()
