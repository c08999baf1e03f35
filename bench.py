#!/usr/bin/env python3
"""Benchmark of the k-max-truss hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload chung_lu_50m] [--plain-only]
    python bench.py --mode batch [--steps K] [--warmup W] [--autotune N]
    python bench.py --mode serve [--queries Q] [--slots S] [--load F]
    python bench.py --mode ingest [--workload rmat22] [--steps K]

A step is one full solve -- CSR build, vertex peel, both passes of support
counting and edge peeling -- from an undirected edge list already resident in
HBM to (kmax, |T|) on the host. The default workload is the largest
single-GPU one: a Chung-Lu power-law graph, gamma 2.1, n = 50,000,000,
505M endpoint samples (500,513,874 edges after deduplication), a planted
60-clique, relabelled, seed 2; the CSR offsets are uint64. `--workload
rmat22` is RMAT scale 22, edge factor 16, (.57,.19,.19,.05), seed 1.

Prints one JSON line on stdout (everything else goes to stderr):
  value        edges/s = m / wall time per solve
  roofline     of the vertex-peel kernels (k_core_push, k_core_cascade) and
               the bit-matrix edge-peel round (k_dense_peel_round), the one with more device
               time per solve: its algorithmic bytes per launch (KERNEL_MODELS,
               the path's per-unit byte figures times the units the launches
               process) over its average launch duration, measured with HIP
               events stamped at each launch's start and end on the solver's
               stream in an extra untimed solve, against 8 TB/s HBM3E;
               `traffic` is the rocprofv3-measured HBM bytes per launch when a
               summary for this kernel and workload is committed under
               profiles/; the other kernel is config.second_kernel_roofline
  config.solve_bytes
               the whole solve's algorithmic bytes (the path's byte model)
               over the wall time, and the rocprofv3-measured HBM bytes of one
               solve over the same time when profiles/ holds them
  config.count_kernel
               the pass-2 support count: for a bit-matrix core, the bytes its
               rows are read with (one source row per workgroup plus one
               destination row per oriented edge, 8 B words, plus 4 B per
               support written) over its event-timed duration
  config.batch, config.serve
               the 4096-graph batched solve (graphs/s against sequential plain
               solves) and the serving loop (closed-loop capacity, p50/p99
               latency under Poisson arrivals); --plain-only skips them
  cpu_baseline the reference's kc_cpu (compiled from its sources into
               oracle/_ref) on a bounded sample of the workload written as
               TSV, timed like the reference's own "All time cost - Reading
               file Time Cost"; the oracle restatement when kc_cpu is missing
               or fails twice

--mode batch / serve print the batch or serving line alone (value in
graphs/s); --mode ingest times the device-side parse of the workload written
as the reference's edge-list text (value in text bytes/s).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import tempfile
import threading
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "kmax-truss wall time & edges/s (1 GPU); batched graphs/s; p99 serve latency"
PROFILES = ROOT / "profiles"
COUNT_KERNELS = {0: "k_support_bitmap", 1: "k_support_edges", 2: "k_dense_support"}
# the GPU box's CPU share per GPU: worker pools there are sized to it
BOX_CPU_SHARE = 16

GENERATORS = {
    "chung_lu_50m": "Chung-Lu gamma 2.1, n 50M, 505M samples, dedup, planted 60-clique, relabelled, seed 2",
    "rmat22": "RMAT(.57,.19,.19,.05) scale 22 ef 16, symmetrised, dedup, relabelled, seed 1",
}
# bounded samples of each workload for the CPU baseline (10-30 s of kc_cpu)
CPU_SAMPLES = {
    "chung_lu_50m": ("chung_lu n=10M, 101M samples (one fifth of the workload's vertices and samples; "
                     "same generator, gamma 2.1, 60-clique, seed 2)",
                     lambda gen: gen.chung_lu(10_000_000, 2.1, 101_000_000, 60, 2)),
    "rmat22": ("the full workload", lambda gen: gen.workload("rmat22")),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def frac_of(achieved_gbs: float, what: str) -> float:
    f = achieved_gbs / HBM_PEAK_GBS
    if f > 1.0:
        raise RuntimeError(f"{what}: {achieved_gbs:.0f} GB/s exceeds the 8 TB/s HBM peak; "
                           "the byte model does not describe this kernel")
    return f


def byte_model(n: int, m: int, st) -> dict:
    """Algorithmic bytes of one solve (the per-stage model of the path):
    build 16m + 8(n+1), vertex peel 8m + 12n + 8, count 16 m_C + 4 W(C+),
    peel 4 W(C+ minus T+)."""
    build = 16 * m + 8 * (n + 1)
    kcore = 8 * m + 12 * n + 8
    count = 16 * st.m_core + 4 * st.w_core
    peel = 4 * st.w_peeled
    return dict(build=build, kcore=kcore, count=count, peel=peel,
                total=build + kcore + count + peel)


def cpu_info() -> dict:
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"host_cpus": len(os.sched_getaffinity(0)), "nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(name: str, budget_s: float) -> dict | None:
    """kc_cpu on a bounded sample of the workload, with the GPU box's CPU
    share of threads; the oracle restatement if kc_cpu is unavailable."""
    from kc_cuda_amd import gen
    from oracle import oracle as O
    what, make = CPU_SAMPLES.get(name, ("the full workload", lambda g: g.workload(name)))
    t0 = time.perf_counter()
    pairs, n = make(gen)
    m = int(len(pairs))
    log(f"[bench] cpu sample ({what}): n={n} m={m} generated in {time.perf_counter() - t0:.1f}s")
    info = cpu_info()
    threads = max(1, min(BOX_CPU_SHARE, info["host_cpus"]))
    if O.kc_cpu_available():
        with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as td:
            tsv = Path(td) / f"{name}_sample.tsv"
            t0 = time.perf_counter()
            gen.write_tsv(tsv, pairs)
            log(f"[bench] wrote {tsv.stat().st_size / 1e9:.2f} GB TSV in {time.perf_counter() - t0:.1f}s")
            for attempt in range(2):
                try:
                    r = O.run_kc_cpu(tsv, threads=threads, timeout=budget_s)
                except Exception as e:  # kc_cpu's parser/peel hazards: retry once
                    log(f"[bench] kc_cpu attempt {attempt} failed: {type(e).__name__}: {e}")
                    continue
                if r.edge_num != m:
                    log(f"[bench] kc_cpu parsed {r.edge_num} edges, expected {m}; discarded")
                    continue
                log(f"[bench] kc_cpu: kmax={r.kmax} |T|={r.truss_edges} t_solve={r.t_solve:.3f}s")
                return dict(value=m / r.t_solve, unit="edges/s", cores=r.threads, kind="reference",
                            sample=f"{what}: m={m} as TSV; t = All time - Reading file "
                                   f"({r.t_solve:.2f} s, kmax {r.kmax}, |T| {r.truss_edges})",
                            seconds=r.t_solve, **info, threads_note=f"{BOX_CPU_SHARE} = the GPU box's CPU "
                            "share per GPU (worker pools there are sized to it)")
    t0 = time.perf_counter()
    s = O.solve_pairs(pairs, n)
    dt = time.perf_counter() - t0
    return dict(value=m / dt, unit="edges/s", cores=threads, kind="port",
                sample=f"{what}: m={m} through the oracle restatement (kmax {s.kmax}, |T| {s.truss_edges})",
                seconds=dt, **info)


def _traffic(kernel: str, workload: str) -> dict | None:
    """A committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE,
    calibrated against a copy of known size), or None."""
    tfile = PROFILES / f"{kernel}_{workload}_traffic.json"
    if not tfile.exists():
        return None
    try:
        return json.loads(tfile.read_text())
    except (OSError, ValueError):
        return None


KERNEL_MODELS = {
    # kernel probe -> (kernel name, bytes from the counters, per-unit figure)
    # levels of up to 2^24 row slots are peeled by k_core_cascade, larger ones by k_core_push rounds
    "vertex_peel": ("k_core_push+k_core_cascade", lambda slots, items: 4 * slots + 12 * items,
                    "4 B per row slot walked + 12 B per vertex taken (the vertex-peel term 8m + 12n + 8)"),
    "dense_peel": ("k_dense_peel_round", lambda slots, items: 4 * slots,
                   "4 B x (d+(u) + d+(v)) per frontier edge walked (the peel term 4 W(C+ minus T+))"),
}


def kernel_roofline(ctx, d, m: int, n: int, workload: str, probe: str) -> dict:
    """One untimed solve with HIP events stamped at every launch's start and
    end of the probed kernel on the solver's stream, and one that counts the
    units the launches process (their atomics stay out of the timing)."""
    name, model, per_unit = KERNEL_MODELS[probe]
    ctx.kernel_probe(probe, 1)
    ctx.solve_device(d, m, n)
    ps = ctx.push_stats()
    ctx.kernel_probe(probe, 2)
    ctx.solve_device(d, m, n)
    counted = ctx.push_stats()
    ctx.kernel_probe(probe, 0)
    launches = max(int(ps["launches"]), 1)
    per_launch = model(counted["slots"], counted["vertices"]) / launches
    us = ps["ms"] * 1e3 / launches
    achieved = per_launch / (us * 1e-6) / 1e9 if us > 0 else 0.0
    log(f"[bench] {name}: {ps['launches']} launches, {us:.1f} us avg ({ps['ms']:.2f} ms in all), "
        f"units {counted['slots']} / {counted['vertices']} -> {achieved:.1f} GB/s")
    t = _traffic(name, workload)
    return {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": frac_of(achieved, name), "traffic": t.get("hbm_bytes_per_launch") if t else None,
            "algorithmic_bytes": per_launch, "launches": int(ps["launches"]), "avg_launch_us": us,
            "total_ms": ps["ms"], "per_unit": per_unit}


def plain_line(args) -> dict:
    from kc_cuda_amd import Context, gen

    t0 = time.perf_counter()
    pairs, n = gen.workload(args.workload)
    m = int(len(pairs))
    log(f"[bench] {args.workload}: n={n} m={m} generated in {time.perf_counter() - t0:.1f}s")

    ctx = Context(0)
    d = ctx.to_device(pairs)
    del pairs
    for i in range(args.warmup):
        r = ctx.solve_device(d, m, n)
        log(f"[bench] warmup {i}: kmax={r.kmax} |T|={r.truss_edges} {r.ms_total:.2f} ms")
    ctx.synchronize()
    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        results.append(ctx.solve_device(d, m, n))
    ctx.synchronize()
    wall = (time.perf_counter() - t0) / max(args.steps, 1)
    r = results[-1]
    for x in results:
        if (x.kmax, x.truss_edges) != (r.kmax, r.truss_edges):
            raise RuntimeError("non-deterministic result")
    st = ctx.stats()
    stages = {k: round(statistics.mean(getattr(x, k) for x in results), 4)
              for k in ("ms_total", "ms_build", "ms_kcore", "ms_pass1", "ms_extract", "ms_count",
                        "ms_peel", "ms_count_kernel")}
    log(f"[bench] timed {args.steps} steps: {wall * 1e3:.2f} ms/step; stages {stages}")
    model = byte_model(n, m, st)
    solve = {"model": "16m + 8(n+1) + 8m + 12n + 8 + 16 m_C + 4 W(C+) + 4 W(C+ minus T+)",
             "model_bytes": model["total"], "model_terms": model,
             "model_frac": frac_of(model["total"] / wall / 1e9, "solve byte model")}
    measured = _traffic("solve", args.workload)
    if measured:
        solve["measured_hbm_bytes"] = measured.get("hbm_bytes_per_solve")
        solve["measured_frac"] = frac_of(solve["measured_hbm_bytes"] / wall / 1e9, "measured solve bytes")
    # the pass-2 support count
    kernel = COUNT_KERNELS.get(st.count_kernel, "k_support_edges")
    count_s = stages["ms_count_kernel"] / 1e3
    if st.count_kernel == 2:
        words = st.core_words
        cbytes = (st.n_core + st.m_core) * words * 8 + 4 * st.m_core
        cmodel = "(n_C + m_C) * ws * 8 B row words + 4 B per support (read from L2/MALL, not HBM)"
    else:
        cbytes = 4 * st.w_core + 4 * st.m_core + 4 * (st.n_core + 1)
        cmodel = "4 W(C+) + 4 m_C + 4 (n_C + 1) (sorted-row merges)"
    count = {"kernel": kernel, "bytes": cbytes, "model": cmodel, "ms": stages["ms_count_kernel"],
             "achieved_gbs": cbytes / count_s / 1e9 if count_s > 0 else 0.0}
    ct = _traffic(kernel, args.workload)
    if ct and count_s > 0:
        count["measured_hbm_bytes"] = ct.get("hbm_bytes_per_launch")
        count["measured_hbm_gbs"] = count["measured_hbm_bytes"] / count_s / 1e9
    roof, other = None, None
    if not args.no_probe:
        # the two largest kernels of a solve; the one with more device time is the line's roofline
        rv = kernel_roofline(ctx, d, m, n, args.workload, "vertex_peel")
        rd = kernel_roofline(ctx, d, m, n, args.workload, "dense_peel") if st.count_kernel == 2 else None
        roof, other = (rd, rv) if rd and rd["total_ms"] > rv["total_ms"] else (rv, rd)
    if args.calibrate_copy_mib > 0:
        nb = args.calibrate_copy_mib << 20
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy_device(b, a, nb)
        log(f"[bench] calibration copy: {nb} bytes device to device")
        a.free()
        b.free()
    d.free()
    ctx.close()

    base = None if args.no_cpu_baseline else cpu_baseline(args.workload, args.cpu_budget)
    return {
        "metric": METRIC, "value": m / wall, "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "m": m, "kmax": r.kmax,
                   "truss_edges": int(r.truss_edges), "generator": GENERATORS.get(args.workload, args.workload),
                   "k_hat": r.k_hat, "max_core": r.max_core, "core_k": r.core_k,
                   "n_core": st.n_core, "m_core": int(st.m_core), "w_core": int(st.w_core),
                   "w_peeled": int(st.w_peeled), "stages_ms": stages, "compactions": r.compactions,
                   "recounts": r.recounts, "solve_bytes": solve, "count_kernel": count,
                   "second_kernel_roofline": other},
        "roofline": roof,
        "cpu_baseline": base,
    }


def batch_summary(args) -> dict:
    """The 4096-graph mix, device-resident, solved as one batch (the union of
    its graphs through one truss decomposition) against sequential plain
    solves of a size-stratified sample. Before the warmup the autotuner
    picks the solver settings on every --autotune-th graph (untimed)."""
    import numpy as np
    from kc_cuda_amd import Context, DeviceArray, gen

    t0 = time.perf_counter()
    pairs, offs, ns = gen.batch()
    count = int(len(ns))
    total = int(offs[-1])
    log(f"[bench] batch: {count} graphs, {total} edges, generated in {time.perf_counter() - t0:.1f}s")
    ctx = Context(0)
    d = ctx.to_device(pairs)
    run = ctx.batch(d, offs, ns)
    tuning = None
    if args.autotune:
        t1 = time.perf_counter()
        _, tuning = ctx.batch_autotune(d, offs, ns, sample_every=args.autotune)
        log(f"[bench] autotune on every {args.autotune}th graph: {tuning} ({time.perf_counter() - t1:.1f}s)")
    steps, warmup = args.batch_steps, args.batch_warmup
    for i in range(warmup):
        t1 = time.perf_counter()
        run.solve()
        log(f"[bench] batch warmup {i}: {time.perf_counter() - t1:.3f}s")
    t1 = time.perf_counter()
    for _ in range(steps):
        run.solve()
    wall = (time.perf_counter() - t1) / max(steps, 1)
    res = run.result()
    if (res.status != 0).any():
        raise RuntimeError("batch reported invalid graphs")
    log(f"[bench] batch: {wall * 1e3:.1f} ms per batch, {count / wall:.0f} graphs/s")
    ctx.tune(0, 0, False)  # the plain solves run with the default settings
    by_size = np.argsort(-(offs[1:] - offs[:-1]).astype(np.int64), kind="stable")
    sample = by_size[::32]
    t_seq = 0.0
    for g in sample:
        mg = int(offs[g + 1] - offs[g])
        view = DeviceArray(ctx, d.ptr + 8 * int(offs[g]), 8 * mg)
        t2 = time.perf_counter()
        r = ctx.solve_device(view, mg, int(ns[g]))
        t_seq += time.perf_counter() - t2
        if (r.kmax, r.truss_edges) != (int(res.kmax[g]), int(res.edges[g])):
            raise RuntimeError(f"graph {g}: plain {r.kmax},{r.truss_edges} batch {res.kmax[g]},{res.edges[g]}")
    seq_gps = len(sample) / t_seq
    log(f"[bench] sequential plain: {len(sample)} sampled graphs in {t_seq:.2f}s -> {seq_gps:.1f} graphs/s")
    run.free()
    d.free()
    ctx.close()
    return {"workload": "batch4096", "graphs": count, "edges": total, "steps": steps, "warmup": warmup,
            "generator": "n log-uniform [64, 65536], ER or R-MAT(.57,.19,.19,.05) 50/50, avg degree U[8,32], "
                         "clique U[4,40], seed 3",
            "engine": "the batch in 4 contiguous parts of about equal edges, each one truss decomposition "
                      "of its union on its own stream, run concurrently; inputs in HBM",
            "ms_per_batch": wall * 1e3, "graphs_per_s": count / wall, "edges_per_s": total / wall,
            "sequential_plain_graphs_per_s": seq_gps,
            "sequential_sample": f"{len(sample)} graphs, every 32nd by size, device-resident",
            "speedup_vs_sequential": (count / wall) / seq_gps, "autotune": tuning}


def _pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else 0.0


def serve_summary(args) -> dict:
    """Serving queries (lognormal sizes, median --median edges, sigma
    --sigma, cap 50M; R-MAT or Chung-Lu with planted cliques of 8-64, seed 4)
    from host memory: the sequential rate (one stand-alone solve after
    another on one context), closed-loop capacity with --slots slots (every
    answer must equal its stand-alone one), then Poisson arrivals at --load
    times that capacity for p50/p99 latency."""
    import numpy as np
    from kc_cuda_amd import Context, Server, gen

    t0 = time.perf_counter()
    mix = gen.serve_mix(args.queries, args.median, args.sigma)
    edges = sum(len(p) for p, _ in mix)
    sizes = np.array([len(p) for p, _ in mix])
    log(f"[bench] serve: {len(mix)} queries, {edges} edges (median {int(np.median(sizes))}, "
        f"max {sizes.max()}), generated in {time.perf_counter() - t0:.1f}s")

    def closed_loop(srv):
        t1 = time.perf_counter()
        tickets = [srv.submit(p, n) for p, n in mix]
        res = [srv.wait(t) for t in tickets]
        return time.perf_counter() - t1, res

    with Context(0) as ctx:
        for p, n in mix:  # warm: buffers grow to the largest query
            ctx.solve_pairs(p, n)
        t1 = time.perf_counter()
        alone = [ctx.solve_pairs(p, n) for p, n in mix]
        t_one = time.perf_counter() - t1
    log(f"[bench] sequential: {len(mix) / t_one:.1f} queries/s")
    with Server(0, args.slots, args.budget_gib << 30) as srv:
        closed_loop(srv)
        c0 = srv.counters()
        t_many, res_many = closed_loop(srv)
        c1 = srv.counters()
        cap = len(mix) / t_many
        log(f"[bench] {args.slots} slots: {cap:.1f} queries/s closed loop, "
            f"{c1['grouped'] - c0['grouped']} of {len(mix)} grouped")
        for i, (a, b) in enumerate(zip(alone, res_many)):
            if b.status or (a.kmax, a.truss_edges) != (b.kmax, b.truss_edges):
                raise RuntimeError(f"query {i}: alone {a.kmax},{a.truss_edges} in the mix "
                                   f"{b.kmax},{b.truss_edges} ({b.error})")
        rate = args.load * cap
        rng = np.random.default_rng(4)
        gaps = rng.exponential(1.0 / rate, len(mix))
        tickets = [0] * len(mix)

        def feed():
            start = time.perf_counter()
            at = 0.0
            for i, (p, n) in enumerate(mix):
                at += gaps[i]
                delay = start + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets[i] = srv.submit(p, n)

        t1 = time.perf_counter()
        th = threading.Thread(target=feed)
        th.start()
        th.join()
        res = [srv.wait(t) for t in tickets]
        wall = time.perf_counter() - t1
        counters = srv.counters()
        for i, (a, b) in enumerate(zip(alone, res)):
            if b.status or (a.kmax, a.truss_edges) != (b.kmax, b.truss_edges):
                raise RuntimeError(f"query {i} (Poisson): alone {a.kmax},{a.truss_edges} in the mix "
                                   f"{b.kmax},{b.truss_edges} ({b.error})")
    lat = [r.total_ms for r in res]
    run = [r.run_ms for r in res]
    log(f"[bench] Poisson at {rate:.1f}/s: p50 {_pct(lat, 50):.1f} ms p99 {_pct(lat, 99):.1f} ms")
    return {"workload": "serve", "queries": len(mix), "edges": int(edges),
            "generator": f"lognormal edges (median {args.median}, sigma {args.sigma}, cap 50M), R-MAT or "
                         "Chung-Lu 50/50, clique U[8,64], seed 4; Poisson arrivals",
            "slots": args.slots, "hbm_budget_gib": args.budget_gib,
            "graphs_per_s": len(mix) / wall, "wall_s": wall,
            "arrival_rate_per_s": rate, "load": args.load,
            "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
            "p50_run_ms": _pct(run, 50), "p99_run_ms": _pct(run, 99),
            "capacity_graphs_per_s": cap, "sequential_graphs_per_s": len(mix) / t_one,
            "capacity_vs_sequential": t_one / t_many,
            "grouped_closed_loop": c1["grouped"] - c0["grouped"],
            "grouped_poisson": counters["grouped"] - c1["grouped"],
            "peak_in_flight": counters["peak_in_flight"],
            "peak_gib_charged": counters["peak_bytes_in_use"] / 2**30,
            "over_estimate": counters["over_estimate"]}


def ingest_line(args) -> dict:
    """The workload written as the reference's edge-list text, read into HBM
    and parsed there (file -> pinned staging -> HBM -> line arrays), checked
    once against the host parser; then --steps solves from the file
    (kt_solve_tsv) after --warmup, each timing its parse (file to line
    arrays in HBM) on its own."""
    import numpy as np
    from kc_cuda_amd import Context, gen, parse_tsv

    pairs, n = gen.workload(args.workload)
    m = int(len(pairs))
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as td:
        tsv = Path(td) / f"{args.workload}.tsv"
        nbytes = gen.write_tsv(tsv, pairs)
        del pairs
        log(f"[bench] ingest: {tsv.stat().st_size / 1e9:.2f} GB text, {2 * m} lines")
        with Context(0) as ctx:
            xs, ys, nn, on_dev = ctx.parse_tsv_device(tsv)
            if not on_dev or nn != n or len(xs) != 2 * m:
                raise RuntimeError("device parse did not take the file")
            t1 = time.perf_counter()
            hx, hy, hn = parse_tsv(tsv)
            t_host = time.perf_counter() - t1
            if hn != nn or not (np.array_equal(hx, xs) and np.array_equal(hy, ys)):
                raise RuntimeError("device and host parsers disagree")
            del hx, hy, xs, ys
            for _ in range(args.warmup):
                ctx.solve_tsv(tsv)
            times, builds, solves = [], [], []
            for _ in range(args.steps):
                r, parse_ms = ctx.solve_tsv(tsv)
                times.append(parse_ms / 1e3)
                builds.append(r.ms_build)
                solves.append(r.ms_total)
            # the same solve with the bucketed CSR build the lines would take
            # if they were not sorted by row
            ctx.tune(unsorted_build=True)
            try:
                rb, _ = ctx.solve_tsv(tsv)
            finally:
                ctx.tune()
            if (rb.kmax, rb.truss_edges) != (r.kmax, r.truss_edges):
                raise RuntimeError("bucketed and run builds disagree")
    t = statistics.median(times)
    log(f"[bench] ingest: file to line arrays in HBM {t * 1e3:.1f} ms (host parser {t_host * 1e3:.0f} ms); "
        f"kmax={r.kmax} |T|={r.truss_edges}")
    return {"metric": "edge-list text to line arrays in HBM (bytes/s)", "value": nbytes / t,
            "unit": "bytes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
            "dtype": "u8", "data": "synthetic",
            "config": {"workload": args.workload, "text_bytes": nbytes, "lines": 2 * m,
                       "timed": "file read, H2D copy and device parse, to line arrays in HBM",
                       "host_parser_ms": t_host * 1e3,
                       "solve_ms": statistics.median(solves),
                       "csr_build_ms": statistics.median(builds),
                       "csr_build_ms_bucketed": rb.ms_build,
                       "csr_build": "lines sorted by row: offsets from their runs, columns copied",
                       "kmax": r.kmax, "truss_edges": int(r.truss_edges)},
            "roofline": None, "cpu_baseline": None}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=None, help="chung_lu_50m (plain default), rmat22 (ingest default)")
    ap.add_argument("--mode", choices=("plain", "batch", "serve", "ingest"), default="plain")
    ap.add_argument("--plain-only", action="store_true", help="plain mode: skip the batch and serve summaries")
    ap.add_argument("--slots", type=int, default=8, help="serving slots")
    ap.add_argument("--budget-gib", type=int, default=96, help="serving HBM budget")
    ap.add_argument("--queries", type=int, default=160, help="serving: queries in the mix")
    ap.add_argument("--median", type=int, default=200_000, help="serving: median edges per query")
    ap.add_argument("--sigma", type=float, default=1.5, help="serving: lognormal sigma of the sizes")
    ap.add_argument("--load", type=float, default=0.7,
                    help="serving: Poisson rate as a fraction of the closed-loop capacity")
    ap.add_argument("--autotune", type=int, default=8,
                    help="batch: autotune on every N-th graph before timing (0: off)")
    ap.add_argument("--batch-steps", type=int, default=None, help="batch: timed batch solves")
    ap.add_argument("--batch-warmup", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the roofline kernel's extra probe solves (under a profiler)")
    ap.add_argument("--cpu-budget", type=float, default=120.0)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after timing, one device-to-device copy of this size (HBM counter "
                         "calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the k-max-truss path here runs on a single MI355X "
              "(multi-GPU sharding is a design note only); rerun with --gpus 1", file=sys.stderr)
        return 2
    if args.workload is None:
        args.workload = "rmat22" if args.mode == "ingest" else "chung_lu_50m"

    if args.mode == "batch":
        args.batch_steps = args.steps if args.batch_steps is None else args.batch_steps
        args.batch_warmup = args.warmup if args.batch_warmup is None else args.batch_warmup
        s = batch_summary(args)
        line = {"metric": METRIC, "value": s["graphs_per_s"], "unit": "graphs/s", "n_gpus": 1,
                "steps": s["steps"], "warmup": s["warmup"], "ms_per_step": s["ms_per_batch"],
                "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "int32",
                "data": "synthetic", "config": s, "roofline": None, "cpu_baseline": None}
    elif args.mode == "serve":
        s = serve_summary(args)
        line = {"metric": METRIC, "value": s["graphs_per_s"], "unit": "graphs/s", "n_gpus": 1, "steps": 1,
                "warmup": 1, "ms_per_step": s["wall_s"] * 1e3, "higher_is_better": True, "scaling": "strong",
                "vs_baseline": None, "dtype": "int32", "data": "synthetic", "config": s, "roofline": None,
                "cpu_baseline": None}
    elif args.mode == "ingest":
        line = ingest_line(args)
    else:
        line = plain_line(args)
        if not args.plain_only:
            args.batch_steps = 3 if args.batch_steps is None else args.batch_steps
            args.batch_warmup = 1 if args.batch_warmup is None else args.batch_warmup
            line["config"]["batch"] = batch_summary(args)
            line["config"]["serve"] = serve_summary(args)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
