#!/usr/bin/env python3
"""MST throughput and batched serving on one MI355X.

Headline workload: GNM n=2^24 vertices, m=2^28 undirected edges (2^29
directed records), uint32 weights uniform in [1, 254], seed 2, generated on
the device by the library's deterministic GNM generator.  One step = one
device-resident solve (input already in HBM, MST left in HBM): the window of
the reference's `mst_complete` timer (benchmarks/mst.cpp:212-230).
value = input edges/s = m / T.

The same JSON line carries the second half of the metric, batched serving
(config.serving: 4096 GNM requests with n log-uniform in [2^8, 2^18], average
degree U[8, 32], Poisson arrivals; batched vs one graph per launch, inputs in
HBM and in pinned host memory), the CPU baselines (cpu_baseline, with every
run in cpu_baseline.runs) and the roofline of the dominant kernel.
"""
import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--log-n", type=int, default=None,
                   help="vertices 2^log_n (default: gnm 24, rgg2d 24, rhg 23)")
    p.add_argument("--log-m", type=int, default=28)
    p.add_argument("--seed", type=int, default=None,
                   help="generator seed (default: gnm 2, rgg2d 3, rhg 4)")
    p.add_argument("--algorithm", choices=["boruvka", "filter_boruvka", "best"], default="best")
    p.add_argument("--filter-threshold", type=int, default=4,
                   help="starting point of the autotuner (or the threshold with --no-autotune)")
    p.add_argument("--no-autotune", action="store_true",
                   help="skip kmst_autotune (filter threshold and sort launch shapes) before "
                        "the timed region")
    p.add_argument("--persistent-max", type=int, default=None,
                   help="rounds on at most this many records run in one persistent launch "
                        "(kmst_tuning.persistent_max; 0 = never; default: the library's)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    p.add_argument("--cpu-threads", type=int, default=None,
                   help="threads of the CPU baseline on the full input (default: "
                        "OMP_NUM_THREADS, else the CPUs this process may run on)")
    p.add_argument("--cpu-c2-4threads", action="store_true",
                   help="also time the full GNM input at 1 x 4 threads (minutes of CPU)")
    p.add_argument("--copy-calibration-mib", type=int, default=0,
                   help="first run one device-to-device copy of this many MiB (a kernel of "
                        "known traffic, to scale rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    p.add_argument("--graph", choices=["gnm", "rgg2d", "rhg"], default="gnm",
                   help="gnm: device GNM (headline); rgg2d: n=2^log_n points, degree 16, weights "
                        "U[1,2^31-1]; rhg: n=2^log_n, degree 32, gamma 2.8 (all generated on "
                        "the device)")
    p.add_argument("--skip-identity", action="store_true",
                   help="skip the plain-vs-filter identity solve (profiling runs: only the "
                        "timed algorithm's launches)")
    p.add_argument("--serving", action="store_true",
                   help="only the batched-serving benchmark (its own JSON line)")
    p.add_argument("--no-serving", action="store_true",
                   help="leave the serving figures out of the throughput line")
    p.add_argument("--requests", type=int, default=4096)
    p.add_argument("--max-batch-records", type=int, default=None,
                   help="serving: admission bound per batch in directed records (default: the "
                        "server's, 2^28)")
    p.add_argument("--rate", type=float, default=20000.0, help="Poisson arrival rate, req/s")
    p.add_argument("--device-inputs", action="store_true",
                   help="--serving: request graphs resident in HBM instead of pinned host memory")
    return p.parse_args()


def cpu_threads(args):
    if args.cpu_threads:
        return args.cpu_threads
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


# ------------------------------------------------------------------ serving
def serving_requests(args, ctx):
    """The request stream: specs (n, m) and every request's arrays generated
    on the device into one pool (3 x total int32)."""
    import numpy as np
    import torch

    rng = np.random.default_rng(args.seed if args.seed is not None else 2)
    specs = []
    for _ in range(args.requests):
        n = int(round(2 ** rng.uniform(8, 18)))
        deg = rng.uniform(8, 32)
        specs.append((n, int(min(n * deg / 2, n * (n - 1) // 2))))
    offs = [0] + [int(x) for x in np.cumsum([2 * m for _, m in specs])]
    total = offs[-1]
    t0 = time.time()
    pool = torch.empty((3, total), dtype=torch.int32, device="cuda")
    base_seed = (args.seed if args.seed is not None else 2) * 1000003
    for i, (n, m) in enumerate(specs):
        a, b = int(offs[i]), int(offs[i + 1])
        ctx.gen_gnm_device(n, m, 254, base_seed + i, pool[0, a:b], pool[1, a:b], pool[2, a:b])
    torch.cuda.synchronize()
    log(f"[serve] {len(specs)} requests, {total} directed records, generated on the device in "
        f"{time.time() - t0:.1f}s")
    return specs, offs, pool, rng


def serve_stream(args, reqs, specs, gaps, max_batch_requests):
    """Submits every request at its Poisson arrival time through a server
    with this batch bound; returns the figures and the per-request results."""
    import threading
    import numpy as np
    import torch
    import kamsta_amd as k

    tuning = None
    if args.persistent_max is not None:
        tuning = dataclasses.replace(k.Tuning(), persistent_max=args.persistent_max)
    srv = k.Server(algorithm=k.FILTER_BORUVKA, filter_threshold=args.filter_threshold,
                   max_batch_requests=max_batch_requests, tuning=tuning,
                   max_batch_records=args.max_batch_records)
    out_offs = np.concatenate([[0], np.cumsum([max(n, 1) for n, _ in specs])])
    try:
        outp = torch.empty((3, int(out_offs[-1])), dtype=torch.int32, pin_memory=True)
    except RuntimeError:
        outp = torch.zeros((3, int(out_offs[-1])), dtype=torch.int32)
    outp = outp.numpy().view(np.uint32)
    outs = [tuple(outp[j, out_offs[i]:out_offs[i + 1]] for j in range(3))
            for i in range(len(reqs))]
    try:
        # warm the server's buffers and kernels on a few requests
        for t in [srv.submit(*reqs[i], out=outs[i]) for i in range(min(8, len(reqs)))]:
            srv.wait(t)
        tickets = [None] * len(reqs)
        arrivals = np.zeros(len(reqs), dtype=np.int64)
        sched = time.monotonic_ns() + 2_000_000 + np.cumsum(gaps * 1e9).astype(np.int64)

        def submitter():
            for i, r in enumerate(reqs):
                while time.monotonic_ns() < sched[i]:
                    pass
                arrivals[i] = sched[i]
                tickets[i] = srv.submit(*r, out=outs[i])

        th = threading.Thread(target=submitter)
        th.start()
        results = []
        for i in range(len(reqs)):
            while tickets[i] is None:
                time.sleep(0.0002)
            results.append(srv.wait(tickets[i]))
        th.join()
        nb, _ = srv.counters()
    finally:
        srv.close()
    done = np.array([r[0]["done_ns"] for r in results], dtype=np.int64)
    lat = (done - arrivals) / 1e6
    span = (done.max() - arrivals.min()) / 1e9
    fig = {"req_per_s": round(len(reqs) / span, 1),
           "p50_ms": round(float(np.percentile(lat, 50)), 3),
           "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "batches": int(nb),
           "largest_batch": int(max(r[0]["batch_requests"] for r in results)),
           "failed": int(sum(r[0]["status"] != 0 for r in results)),
           "span_s": round(span, 3)}
    return fig, results


def serving_mode(args, specs, offs, pool, rng, inputs):
    """Batched and one-graph-per-launch streams on the same arrival trace,
    with the request arrays in HBM (inputs="device") or in pinned host memory
    (inputs="pinned"); results are compared request by request."""
    import numpy as np
    import torch

    host_memory = "device"
    if inputs == "device":
        src = pool
        reqs = [(src[0, offs[i]:offs[i + 1]], src[1, offs[i]:offs[i + 1]],
                 src[2, offs[i]:offs[i + 1]], n) for i, (n, _) in enumerate(specs)]
    else:
        # pinned blocks freed earlier in the run stay cached by torch: return
        # them first.  torch rounds a pinned allocation up to a power of two,
        # so the requests go into pinned chunks of exactly 2 GiB
        if hasattr(torch._C, "_host_emptyCache"):
            torch._C._host_emptyCache()
        chunk = 1 << 29  # int32 entries
        chunks, reqs, used = [], [], chunk
        host_memory = "pinned"
        for i, (n, m) in enumerate(specs):
            need = 6 * m
            if used + need > chunk:
                try:
                    chunks.append(torch.empty(max(chunk, need), dtype=torch.int32,
                                              pin_memory=host_memory == "pinned"))
                except RuntimeError as e:
                    log(f"[serve] pinned allocation failed ({e}); pageable host memory")
                    host_memory = "pageable"
                    chunks.append(torch.empty(max(chunk, need), dtype=torch.int32))
                used = 0
            seg = chunks[-1][used:used + need].view(3, 2 * m)
            seg.copy_(pool[:, offs[i]:offs[i + 1]])
            used += need
            sv = seg.numpy().view(np.uint32)
            reqs.append((sv[0], sv[1], sv[2], n))
        torch.cuda.synchronize()
        log(f"[serve] request arrays in {len(chunks)} {host_memory} chunks of 2 GiB")
    gaps = rng.exponential(1.0 / args.rate, size=len(reqs))
    batched, rb = serve_stream(args, reqs, specs, gaps, len(reqs))
    log(f"[serve] {inputs} inputs, batched: {batched}")
    alone, ra = serve_stream(args, reqs, specs, gaps, 1)
    log(f"[serve] {inputs} inputs, one graph per launch: {alone}")
    identical = all(x[0]["status"] == 0 and x[0]["total_weight"] == y[0]["total_weight"] and
                    x[0]["mst_edges"] == y[0]["mst_edges"] and
                    all(np.array_equal(p, q) for p, q in zip(x[1], y[1]))
                    for x, y in zip(rb, ra))
    return {"inputs": host_memory, "batched": batched, "one_graph_per_launch": alone,
            "speedup_req_per_s": round(batched["req_per_s"] / alone["req_per_s"], 2),
            "identical_results": bool(identical)}


def serving_figures(args, ctx):
    specs, offs, pool, rng = serving_requests(args, ctx)
    modes = []
    for inputs in ("device", "pinned"):
        modes.append(serving_mode(args, specs, offs, pool, rng, inputs))
    del pool
    return {
        "workload": (f"{args.requests} GNM requests, n log-uniform [2^8,2^18], avg degree "
                     f"U[8,32], weights U[1,254], seed base + i, Poisson arrivals "
                     f"{args.rate:g}/s"),
        "directed_records": int(offs[-1]),
        "algorithm": "filter_boruvka",
        "max_batch_records": args.max_batch_records or (1 << 28),
        "latency": "arrival to result on the host (includes upload from host memory when the "
                   "inputs are there, and the download of the MST)",
        "modes": modes,
    }


def serving_only(args):
    import kamsta_amd as k
    if args.seed is None:
        args.seed = 2
    ctx = k.Context(0)
    try:
        specs, offs, pool, rng = serving_requests(args, ctx)
        mode = serving_mode(args, specs, offs, pool, rng,
                            "device" if args.device_inputs else "pinned")
        del pool
    finally:
        ctx.close()
    line = {
        "metric": "batched-serving req/s + p50/p99",
        "value": mode["batched"]["req_per_s"],
        "unit": "requests/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "u32/u64",
        "data": f"synthetic: {args.requests} GNM requests, inputs in {mode['inputs']} memory",
        "config": {"workload": "batched serving", "requests": args.requests, "rate": args.rate,
                   "directed_records": int(offs[-1]), "algorithm": "filter_boruvka"},
        **mode,
    }
    print(json.dumps(line), flush=True)


# ------------------------------------------------------------------ CPU baselines
def cpu_baselines(args, ctx, full, threads):
    """The oracle (the CPU restatement of hybridMST::boruvka) timed on the
    host cores, one process (MPI ranks = 1) x T OpenMP threads, on
    - GNM n=2^14, m=2^18, seed 0x4B414D5354410001, at
      exactly 4 threads, local_preprocessing 1 (the driver default: local
      filter-Boruvka) and 0 (boruvka_core rounds);
    - the headline input at full size with `threads` threads (and, with
      --cpu-c2-4threads, at 4).
    Each GPU result on the same arrays is checked identical."""
    import numpy as np
    import torch
    import kamsta_amd as k
    import oracle

    def gpu_solve(s, d, w, N, n, algorithm=k.FILTER_BORUVKA):
        outs = [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3)]
        cnt, tw = ctx.solve_device(s, d, w, N, n, *outs, n, algorithm=algorithm)
        torch.cuda.synchronize()
        return cnt, tw, [o[:cnt].cpu().numpy().view(np.uint32) for o in outs]

    def same(g, cm, cw):
        cnt, tw, (gs, gd, gw) = g
        return bool(tw == cw and cnt == len(cm) and np.array_equal(gs, cm["src"]) and
                    np.array_equal(gd, cm["dst"]) and np.array_equal(gw, cm["weight"]))

    def timed(rec, level, thr, reps):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            cm, cw = oracle.boruvka(rec, level, thr)
            ts.append(time.perf_counter() - t0)
        ts.sort()
        return ts[len(ts) // 2], cm, cw

    runs = []
    # config 1
    n1, m1 = 1 << 14, 1 << 18
    arr = [torch.empty(2 * m1, dtype=torch.int32, device="cuda") for _ in range(3)]
    ctx.gen_gnm_device(n1, m1, 254, 0x4B414D5354410001, *arr)
    host = [a.cpu().numpy().view(np.uint32) for a in arr]
    rec = oracle.wedges_from_soa(*host)
    g1 = gpu_solve(*arr, 2 * m1, n1, k.BORUVKA)
    for level in (1, 0):
        oracle.boruvka(rec, level, 4)  # warm-up
        t, cm, cw = timed(rec, level, 4, 5)
        runs.append({"config": "GNM n=2^14 m=2^18 weights U[1,254] seed 0x4B414D5354410001",
                     "algorithm": "boruvka, local_preprocessing %d (%s)" % (
                         level, "local filter-Boruvka" if level else "boruvka_core rounds"),
                     "processes": 1, "threads": 4, "seconds": round(t, 4),
                     "value": round(m1 / t, 1), "unit": "input edges/s",
                     "timing": "median of 5 after a warm-up",
                     "gpu_identical": same(g1, cm, cw)})
        log(f"[bench] cpu baseline: {runs[-1]}")
    del arr, host, rec
    # the headline input at full size
    s, d, w, N, n, m = full
    t0 = time.perf_counter()
    host = [a.cpu().numpy().view(np.uint32) for a in (s, d, w)]
    rec = oracle.wedges_from_soa(*host)
    del host
    log(f"[bench] cpu baseline: full input on the host in {time.perf_counter() - t0:.1f}s")
    g2 = gpu_solve(s, d, w, N, n)
    plan = [threads] + ([4] if args.cpu_c2_4threads and threads != 4 else [])
    for thr in plan:
        t, cm, cw = timed(rec, 1, thr, 1)
        runs.append({"config": "GNM n=2^24 m=2^28 weights U[1,254] seed %d" % args.seed,
                     "algorithm": "boruvka, local_preprocessing 1 (local filter-Boruvka; on one "
                                  "process filter_boruvka computes the same)",
                     "processes": 1, "threads": thr, "seconds": round(t, 3),
                     "value": round(m / t, 1), "unit": "input edges/s",
                     "timing": "one run (no warm-up: a run is tens of seconds)",
                     "gpu_identical": same(g2, cm, cw)})
        log(f"[bench] cpu baseline: {runs[-1]}")
    if not args.cpu_c2_4threads:
        runs.append({"config": "GNM n=2^24 m=2^28", "processes": 1, "threads": 4,
                     "skipped": "about four minutes of CPU at 4 threads, beyond the bench's time "
                                "budget; run bench.py --cpu-c2-4threads for it"})
    del rec
    head = runs[2] if len(runs) > 2 else runs[0]
    return {"value": head["value"], "unit": "input edges/s", "cores": head["threads"],
            "kind": "port",
            "sample": (f"{head['config']}, oracle {head['algorithm']}, 1 process x "
                       f"{head['threads']} OpenMP threads, {head['seconds']} s; GPU result "
                       f"identical: {head['gpu_identical']}"),
            "build": "g++ -O3 -march=x86-64-v3 -fopenmp (the reference builds -march=native; "
                     "the oracle library is built once and runs on the host of another machine)",
            "host_cpus": len(os.sched_getaffinity(0)),
            "runs": runs}


# ------------------------------------------------------------------ throughput
def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "the solver runs on one GPU: sharding a solve over several "
                                    "GPUs is designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.serving:
        serving_only(args)
        return
    import numpy as np
    import torch
    import kamsta_amd as k

    if args.log_n is None:
        args.log_n = {"gnm": 24, "rgg2d": 24, "rhg": 23}[args.graph]
    if args.seed is None:
        args.seed = {"gnm": 2, "rgg2d": 3, "rhg": 4}[args.graph]
    n, m = 1 << args.log_n, 1 << args.log_m
    N = 2 * m
    dev = torch.device("cuda:0")
    workload = f"GNM n=2^{args.log_n} m=2^{args.log_m} weights U[1,254]"
    if args.copy_calibration_mib:
        cal = torch.empty(args.copy_calibration_mib << 18, dtype=torch.int32, device=dev)
        cal.fill_(1)
        cal2 = cal.clone()
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.copy_calibration_mib} MiB done")
        del cal, cal2
    ctx = k.Context(0)
    if args.persistent_max is not None:
        ctx.set_tuning(dataclasses.replace(ctx.get_tuning(), persistent_max=args.persistent_max))
    t0 = time.time()

    def generate(seed):
        if args.graph == "gnm":
            a = [torch.empty(N, dtype=torch.int32, device=dev) for _ in range(3)]
            ctx.gen_gnm_device(n, m, 254, seed, *a)
            return a
        if args.graph == "rgg2d":
            return ctx.gen_rgg2d_device(n, 16.0, (1 << 31) - 1, seed)
        return ctx.gen_rhg_device(n, 32.0, 2.8, 254, seed)

    log(f"[bench] generating {args.graph} n=2^{args.log_n}")
    s, d, w = generate(args.seed)
    if args.graph == "rgg2d":
        workload = f"RGG2D n=2^{args.log_n} avg degree 16 weights U[1,2^31-1]"
    elif args.graph == "rhg":
        workload = f"RHG n=2^{args.log_n} avg degree 32 gamma 2.8 weights U[1,254]"
    N = s.numel()
    m = N // 2
    torch.cuda.synchronize()
    log(f"[bench] {workload}: {N} records, generated in {time.time() - t0:.2f}s")
    outs = [torch.empty(n, dtype=torch.int32, device=dev) for _ in range(3)]
    outs_f = [torch.empty(n, dtype=torch.int32, device=dev) for _ in range(3)]
    cfg = k.AlgorithmConfig(filter_threshold=args.filter_threshold)

    def solve(algo, o, c=None):
        return ctx.solve_device(s, d, w, N, n, o[0], o[1], o[2], n, algorithm=algo,
                                config=c or cfg)

    # exactness across algorithms (size-independent: the MSF is unique)
    if args.skip_identity:
        cnt_p, tw_p = None, None
    else:
        cnt_p, tw_p = solve(k.BORUVKA, outs)
        cnt_f, tw_f = solve(k.FILTER_BORUVKA, outs_f)
        torch.cuda.synchronize()
        same = (cnt_p == cnt_f and tw_p == tw_f and all(
            torch.equal(a[:cnt_p], b[:cnt_p]) for a, b in zip(outs, outs_f)))
        log(f"[bench] plain: {cnt_p} edges weight {tw_p}; filter: {cnt_f} edges weight {tw_f}; "
            f"identical={same}")
        if not same:
            print(json.dumps({"error": "plain and filter MSTs differ"}))
            sys.exit(1)

    def time_steps(algo, steps, warmup, c=None):
        for _ in range(warmup):
            solve(algo, outs, c)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            last = solve(algo, outs, c)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / steps, last

    # the library defaults before any tuning (filter threshold 4, default
    # launch shapes), for comparison with the tuned figure
    default_ms = None
    if not args.no_autotune:
        default_ms, _ = time_steps(k.FILTER_BORUVKA, args.steps, args.warmup)
        log(f"[bench] filter_boruvka with the library defaults: {default_ms:.3f} ms/step")

    # autotuner: filter threshold and sort launch shapes, timed on another
    # input of the same shape (seed + 1), not on the timed one; every
    # candidate is checked to return the same MST
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        ts, td, tw_ = generate(args.seed + 1)
        res = ctx.autotune(ts, td, tw_, ts.numel(), n, algorithm=k.FILTER_BORUVKA, config=cfg,
                           repeats=3)
        del ts, td, tw_
        cfg = k.AlgorithmConfig(filter_threshold=res["filter_threshold"])
        tune = {"input": f"seed {args.seed + 1} (same shape as the timed input)",
                "filter_threshold": res["filter_threshold"],
                "sort_threads": res["tuning"].sort_threads,
                "balanced_digits": res["tuning"].balanced_digits,
                "sort_wide_min": res["tuning"].sort_wide_min,
                "persistent_max": res["tuning"].persistent_max,
                "best_ms": round(res["best_ms"], 3), "candidates": res["candidates"],
                "seconds": round(time.perf_counter() - t0, 2),
                "default_ms_per_step": round(default_ms, 3)}
        log(f"[bench] autotune: {tune}")

    def timed(algo, steps, warmup):
        """The timed region runs without per-launch instrumentation; a second
        pass over the same number of steps records HIP events around every
        launch (on the library's stream) for the kernel breakdown and the
        roofline.  The events widen the gaps between kernels, not the
        kernels, so the per-kernel times hold for the first pass too."""
        ms, last = time_steps(algo, steps, warmup)
        ctx.set_timing(True)
        t2 = time.perf_counter()
        for _ in range(steps):
            solve(algo, outs)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        st = ctx.stats()
        ctx.set_timing(False)
        st["last_result"] = last
        st["ms_per_step_instrumented"] = (t3 - t2) * 1000.0 / steps
        return ms, st

    results = {}
    algos = {"boruvka": k.BORUVKA, "filter_boruvka": k.FILTER_BORUVKA}
    chosen = list(algos) if args.algorithm == "best" else [args.algorithm]
    for name in chosen:
        ms, st = timed(algos[name], args.steps, args.warmup)
        results[name] = (ms, st)
        log(f"[bench] {name}: {ms:.2f} ms/step, {m / (ms / 1000.0) / 1e9:.3f} G input edges/s "
            f"({st['ms_per_step_instrumented']:.2f} ms/step with per-launch events)")
        kt = st["kernels"]
        log(f"[bench] {name} kernels: " + json.dumps(
            {kk: {"ms_per_step": round(v["ms"] / args.steps, 3), "launches": v["launches"] / args.steps,
                  "GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)}
             for kk, v in sorted(kt.items(), key=lambda x: -x[1]["ms"])}))
        log(f"[bench] {name} rounds: " + json.dumps(st["rounds"][:40]))
        log(f"[bench] {name} filter: " + json.dumps(st["filter"]))
    best = min(results, key=lambda a: results[a][0])
    ms, st = results[best]
    if cnt_p is None:
        cnt_p, tw_p = st["last_result"]

    # dominant kernel of the timed region: algorithmic bytes / average duration
    kern = st["kernels"]
    dom = max(kern, key=lambda kk: kern[kk]["ms"])
    kd = kern[dom]
    per_launch_bytes = kd["bytes"] / kd["launches"]
    avg_ms = kd["ms"] / kd["launches"]
    achieved = per_launch_bytes / (avg_ms / 1000.0) / 1e9
    # HBM bytes per launch of the same kernel from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs (profiles/traffic_*.md) on the same
    # workload, or null.  The counted runs may split the kind into more or
    # fewer launches than this run's tuning does (the persistent tails take
    # the small rounds), so their ratio of HBM to algorithmic bytes is applied
    # to this run's algorithmic bytes per launch.
    traffic = None
    traffic_ratio = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            tj = json.load(open(tfile))
            t = tj.get(best, {}).get(dom) if tj.get("workload") == workload else None
            if t and t.get("traffic_ratio"):
                traffic_ratio = t["traffic_ratio"]
                traffic = round(traffic_ratio * per_launch_bytes)
            elif t:
                traffic = round(t["bytes_per_launch"])
        except Exception:
            traffic = None
    total_bytes = sum(v["bytes"] for v in kern.values()) / args.steps
    # sort / compaction kernels together (relabel+compact, radix passes,
    # light/heavy split): algorithmic bytes over their summed time
    sc = [kern[x] for x in ("relabel_scatter", "sort_hist", "sort_scatter", "split_scatter",
                            "dedup") if x in kern]
    sc_gbs = sum(v["bytes"] for v in sc) / max(sum(v["ms"] for v in sc), 1e-9) / 1e6
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "traffic_ratio": round(traffic_ratio, 3) if traffic_ratio else None,
                "traffic_basis": ("HBM bytes over algorithmic bytes of this kernel kind in the "
                                  "committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs on this "
                                  "workload (profiles/traffic_*.md), times this run's "
                                  "algorithmic bytes per launch") if traffic_ratio else None,
                "algorithmic_bytes_per_launch": round(per_launch_bytes),
                "avg_launch_ms": round(avg_ms, 4),
                "end_to_end_frac": round(total_bytes / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
                "sort_compaction_frac": round(sc_gbs / HBM_PEAK_GBS, 4),
                "timing": ("HIP events around every launch on the library's stream, in a second "
                           f"pass of {args.steps} steps after the timed region "
                           f"({st['ms_per_step_instrumented']:.3f} ms/step with the events)")}

    # PCIe-inclusive time of the best algorithm: input uploaded from host
    # memory (pinned when the allocation succeeds), solve, MST copied back
    pcie = None
    try:
        pin = True
        try:
            host = [torch.empty(N, dtype=torch.int32, pin_memory=True) for _ in range(3)]
        except RuntimeError:
            pin = False
            host = [torch.empty(N, dtype=torch.int32) for _ in range(3)]
        for h, dv in zip(host, (s, d, w)):
            h.copy_(dv)
        out_h = [torch.empty(n, dtype=torch.int32, pin_memory=pin) for _ in range(3)]
        torch.cuda.synchronize()
        best_t = None
        for _ in range(2):
            t0 = time.perf_counter()
            for h, dv in zip(host, (s, d, w)):
                dv.copy_(h, non_blocking=pin)
            # the library solves on its own stream: the upload must be complete
            torch.cuda.current_stream().synchronize()
            c, _ = solve(algos[best], outs)
            for oh, od in zip(out_h, outs):
                oh[:c].copy_(od[:c], non_blocking=pin)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            best_t = t if best_t is None else min(best_t, t)
        pcie = {"ms": round(best_t * 1000.0, 3), "input_edges_per_s": round(m / best_t, 1),
                "host_memory": "pinned" if pin else "pageable",
                "bytes_h2d": 12 * N, "bytes_d2h": 12 * c, "same_mst_edges": c == cnt_p,
                "note": "upload of the (src, dst, weight) arrays + solve + download of the MST; "
                        "not the headline"}
        log(f"[bench] PCIe-inclusive: {pcie}")
        del host, out_h
    except RuntimeError as e:
        log(f"[bench] PCIe-inclusive run skipped: {e}")

    serving = None
    if not args.no_serving and args.requests > 0:
        serving = serving_figures(args, ctx)

    cpu = None
    if not args.no_cpu and args.graph == "gnm":
        cpu = cpu_baselines(args, ctx, (s, d, w, N, n, m), cpu_threads(args))

    line = {
        "metric": "input edges/s at GNM 2^28 (bit-exact MST)" if args.graph == "gnm"
                  else "input edges/s (bit-exact MST)",
        "value": round(m / (ms / 1000.0), 1),
        "unit": "input edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64",
        "data": "synthetic (device %s generator, seed %d)" % (args.graph.upper(), args.seed),
        "config": {"workload": workload,
                   "algorithm": best, "filter_threshold": cfg.filter_threshold,
                   "persistent_max": ctx.get_tuning().persistent_max,
                   "autotune": tune,
                   "directed_records": N,
                   "mst_edges": cnt_p, "mst_weight": tw_p,
                   "ms_per_step": {a: round(results[a][0], 3) for a in results},
                   "pcie_inclusive": pcie,
                   "serving": serving},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
