#!/usr/bin/env python3
"""Benchmark of the strand-generation hot path on one MI355X.

Workload (one step = one full generation pass, guide preparation + strand
kernel, over one synthetic description resident in HBM):
  plane-524288tri: 512x512-quad displaced plane (524,288 triangles) with
  log-normal(0, 0.5) area weights, 4,096 guides x 16 CVs, stencil K=8,
  2,000,000 strands x 16 CVs, taper 0.01 -> 0.001, two stacked noise octaves
  (0.1 @ 4, 0.05 @ 8); outputs 32,000,000 float4 points + 2,000,000 counts.

Prints one JSON line: points/s, the strand kernel's HBM roofline fraction
(algorithmic bytes / kernel time, timed with HIP events on the launch
stream), the CPU oracle's points/s on a bounded sample of the same workload
on this host's cores, and the metric's surrogate clause: the learned
surrogate's pipeline against the exact expression runtime's on the
1,000,000-root, 24-CV plane workload at the 1e-3 relative threshold
("surrogate": trained outside the timed region).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "output points/sec + % HBM roofline; surrogate vs plain expr at <=1e-3 rel err"
HBM_PEAK_GBS = 8000.0
SIMD_CYCLES_PER_S = 1024 * 2.4e9  # 256 CUs x 4 SIMDs at the 2.4 GHz peak engine clock


def valu_issue(insts, ms, key):
    """A kernel's VALU issue: its profiled wave64 VALU instructions over its
    measured time, and the SIMD cycles between two of them at the peak clock.
    A full-rate FP32 wave64 instruction occupies a SIMD for 4 cycles; a kernel
    near (or, when some of its instructions take fewer, below) 4 issues VALU
    all the time."""
    rate = insts / (ms * 1e-3)
    return {"achieved": round(rate / 1e9, 2), "unit": "G wave-instructions/s",
            "simd_cycles_per_instruction": round(SIMD_CYCLES_PER_S / rate, 2), "full_rate_cycles": 4, key: insts}
PMC_SUMMARY = os.path.join(ROOT, "profiles", "strand_kernel_plane2M_pmc.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """The host's CPU model, its logical CPU count and this process's
    affinity set: the facts a CPU baseline is read against."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def baseline_threads():
    """Worker threads for a CPU baseline: the affinity set, at most 16."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def sampled_seconds(run, N, target_s, first=65536):
    """Time run(s) (seconds for the first s strands) on a sample sized to
    about target_s of work, capped at N; whole-workload runs repeat (median
    of up to 5).  Returns (s, median seconds, runs)."""
    s = min(N, first)
    dt = run(s)
    s = int(min(N, max(s, s * target_s / max(dt, 1e-6))))
    times = [run(s)]
    while s == N and sum(times) < target_s and len(times) < 5:
        times.append(run(s))
    return s, sorted(times)[len(times) // 2], len(times)


def cpu_baseline(blob, w, target_s, threads):
    """The oracle's multithreaded CPU path (128-strand tiles over an atomic
    counter, asset.cpp:160-202) on the same workload: per-root attribute
    programs first where the workload has them, then generation.  The sample
    is the first S strands, S sized to about target_s of work and capped at
    the whole workload; whole-workload runs are repeated (median of up to 5)."""
    import numpy as np
    import oracle
    oracle.lib()
    programs, maps = w.info.get("programs"), w.info.get("maps", [])

    def run(strands):
        t = time.perf_counter()
        attrs = None
        if programs:
            attrs, _ = oracle.root_attributes(blob, w.params, programs, maps, strands=strands, workers=threads)
        oracle.generate(blob, w.params, octaves=w.octaves, strands=strands, workers=threads, attributes=attrs)
        return time.perf_counter() - t

    dt = run(np.arange(min(4096, w.params.strand_count), dtype=np.uint32))
    s = int(min(w.params.strand_count, max(4096, 4096 * target_s / max(dt, 1e-6))))
    sample = np.arange(s, dtype=np.uint32)
    times = [run(sample)]
    while s == w.params.strand_count and sum(times) < target_s and len(times) < 5:
        times.append(run(sample))
    dt = sorted(times)[len(times) // 2]
    pts = s * w.params.cvs_per_strand
    what = "attribute programs + generation" if programs else "generation"
    return dict({"value": pts / dt, "unit": "points/s", "cores": threads, "kind": "port",
                 "sample": f"strands 0..{s - 1} of the same description ({pts} points), {what}, median "
                           f"{dt:.3f} s of {len(times)} run(s); float restatement of the reference CPU path "
                           f"(oracle/), {threads} threads"}, **host_cpu())


SURROGATE_PMC_SUMMARY = os.path.join(ROOT, "profiles", "surrogate_kernel_plane1M_pmc.json")
SURROGATE_FLOP_PER_ROOT = 26368  # 2 x (10x64 + 3 x 64x64 + 64x4) multiply-adds of the network


def surrogate_record(args, w=None, threshold=None, stride=None, epochs=None, steps=None, quiet=False):
    """The learned-surrogate workload (1,000,000 roots x 24 CVs on the plane,
    four 40-op attribute programs): train on 2^20 exact-runtime pairs from
    seed-disjoint roots (outside the timed region), then time two pipelines
    over the same roots, each preparing the description once:
      exact:     prepare -> exact attributes -> strands
      surrogate: prepare -> fused root features + MLP (triangle order) ->
                 exact check of every stride-th root, exact re-evaluation of
                 each 64-root tile that misses the threshold -> strands.
    Returns the record (pipelines, kernel times, served tiles, errors of the
    served roots against the exact runtime, training)."""
    import torch
    import nanoxgen_amd as nx
    from nanoxgen_amd import expr, surrogate, workloads

    w = w or workloads.config4()
    threshold = args.threshold if threshold is None else threshold
    stride = args.sample_stride if stride is None else stride
    epochs = args.train_epochs if epochs is None else epochs
    steps = args.steps if steps is None else steps
    N, Cv = w.params.strand_count, w.params.cvs_per_strand
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    es, maps = expr.ExprSet(w.info["programs"]), expr.TexelMaps(w.info["maps"])
    pairs = 1 << 20
    train_p = w.params.copy(seed=w.info["train_seed"], strand_count=pairs)
    t0 = time.perf_counter()
    feats = surrogate.root_features(asset, train_p, maps)
    targets = torch.empty((pairs, 4), device="cuda")
    expr.evaluate_root_attributes(asset, train_p, es, maps, targets)
    net = surrogate.Surrogate(seed=1)
    net.fit_norm(feats, targets)
    hist = net.train(feats, targets, epochs=epochs, batch=4096, lr=3e-3, log_every=256)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del feats, targets
    if not quiet:
        log(f"trained on {pairs} pairs, {epochs} epochs: loss {hist[0][1]:.4f} -> {hist[-1][1]:.2e} in {train_s:.1f} s")

    out = nx.PackedOutputs(N, Cv, roots=False, root_uvs=False, point_counts=True)
    ws = torch.empty(asset.workspace_bytes(w.params), dtype=torch.uint8, device="cuda")
    attrs = torch.empty((N, 4), dtype=torch.float32, device="cuda")
    flags = torch.zeros((N + 63) // 64, dtype=torch.int32, device="cuda")
    stats = torch.zeros(2, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()

    def exact_attrs():
        expr.evaluate_root_attributes(asset, w.params, es, maps, attrs, stream=stream)

    def surrogate_attrs():
        surrogate.surrogate_attributes(asset, w.params, net, maps, attrs, stream=stream, prepared=ws)

    def verify():
        surrogate.verify_attributes(asset, w.params, es, maps, attrs, threshold, stride, flags, stats, stream=stream)

    def pipeline(fns):
        def step():
            nx.prepare_packed(asset, w.params, ws, stream=stream)
            for fn in fns:
                fn()
            nx.generate_packed_prepared(asset, w.params, out, ws, octaves=w.octaves, stream=stream, attributes=attrs)
        return step

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / steps

    def event_ms(fn, reps=7):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ms = []
        for _ in range(reps):
            ev[0].record(stream)
            fn()
            ev[1].record(stream)
            torch.cuda.synchronize()
            ms.append(ev[0].elapsed_time(ev[1]))
        return sorted(ms)[reps // 2]

    # both pipelines again with consecutive steps overlapped: each step's
    # attributes (and, for the surrogate, its preparation, network and
    # check) on a side stream beside the previous step's strand kernel
    from nanoxgen_amd.pipeline import GenerationPipeline

    def exact_fn(params, out_attrs, side):
        expr.evaluate_root_attributes(asset, params, es, maps, out_attrs, stream=side)

    def surrogate_fn(params, out_attrs, side, prepared):
        surrogate.surrogate_attributes(asset, params, net, maps, out_attrs, stream=side, prepared=prepared)
        surrogate.verify_attributes(asset, params, es, maps, out_attrs, threshold, stride, flags, stats, stream=side)

    pipe_ms = {}
    for name, fn, after in (("exact", exact_fn, False), ("surrogate", surrogate_fn, True)):
        gp = GenerationPipeline(asset, w.params, octaves=w.octaves, stream=stream, attribute_fn=fn,
                                attributes_after_prepare=after)
        pipe_ms[name] = timed(lambda: gp.submit(out))
        del gp
    stats.zero_()
    exact_ms = timed(pipeline([exact_attrs]))
    exact = attrs.clone()
    sur_ms = timed(pipeline([surrogate_attrs, verify]))
    torch.cuda.synchronize()
    tiles_redone, samples = stats.tolist()
    counts_exact = bool((out.point_counts == Cv).all())
    rel = ((attrs - exact).abs() / exact.abs().clamp_min(1e-6)).max(1).values
    served = torch.repeat_interleave(flags == 0, 64)[:N]
    rel_served = rel[served]
    # the attribute phase alone, with the workspace prepared
    nx.prepare_packed(asset, w.params, ws, stream=stream)
    ex_attr_ms = event_ms(exact_attrs)
    mlp_ms = event_ms(surrogate_attrs)
    surrogate_attrs()
    verify_ms = event_ms(lambda: (surrogate_attrs(), verify())) - mlp_ms
    # before the gate: the network's own error over every root
    surrogate_attrs()
    torch.cuda.synchronize()
    rel_net = ((attrs - exact).abs() / exact.abs().clamp_min(1e-6)).max(1).values
    q = torch.tensor([0.5, 0.99], device="cuda")
    points = N * Cv
    rec = {"workload": w.name, "frequency_scale": w.info.get("frequency_scale"), "threshold": threshold,
           "sample_stride": stride,
           "exact_pipeline_points_per_s": points / (exact_ms * 1e-3), "exact_pipeline_ms": round(exact_ms, 4),
           "surrogate_pipeline_points_per_s": points / (sur_ms * 1e-3), "surrogate_pipeline_ms": round(sur_ms, 4),
           "speedup_vs_exact": round(exact_ms / sur_ms, 4),
           "pipelined_steps": {"exact_pipeline_ms": round(pipe_ms["exact"], 4),
                               "surrogate_pipeline_ms": round(pipe_ms["surrogate"], 4),
                               "speedup_vs_exact": round(pipe_ms["exact"] / pipe_ms["surrogate"], 4)},
           "attributes_ms_exact": round(ex_attr_ms, 4), "surrogate_kernel_ms": round(mlp_ms, 4),
           "check_and_fallback_ms": round(verify_ms, 4),
           "tiles": int(flags.numel()), "tiles_served": int(flags.numel()) - tiles_redone,
           "tiles_reevaluated": tiles_redone, "samples_checked": samples,
           "served_roots": int(served.sum().item()),
           "served_roots_within_threshold": float((rel_served <= threshold).float().mean().item())
           if rel_served.numel() else None,
           "served_roots_max_rel_error": float(rel_served.max().item()) if rel_served.numel() else None,
           "network_rel_error_p50_p99": [float(v) for v in torch.quantile(rel_net[:1 << 20], q)],
           "network_roots_within_threshold": float((rel_net <= threshold).float().mean().item()),
           "point_counts_exact": counts_exact,
           "train_pairs": pairs, "train_epochs": epochs, "train_seconds": round(train_s, 2),
           "train_loss_first_last": [hist[0][1], hist[-1][1]]}
    return rec


def bench_surrogate(args):
    """The learned-surrogate workload's own line: value is the surrogate
    pipeline's points/s; the roofline object is the surrogate kernel's
    (FLOPs of the network at the BF16 MFMA peak, with the VALU issue rate of
    its SiLU activations beside it from the profiled counters)."""
    rec = surrogate_record(args)
    N = 1_000_000
    mlp_ms = rec["surrogate_kernel_ms"]
    flops = SURROGATE_FLOP_PER_ROOT * N / (mlp_ms * 1e-3) / 1e12
    valu = None
    if os.path.exists(SURROGATE_PMC_SUMMARY):
        # VALU issue roofline: the profiled wave64 VALU instructions per
        # launch weighted by their issue cycles (4, and 16 for the
        # quarter-rate transcendentals of SiLU), over the kernel time,
        # against 1,024 SIMDs x 2.4 GHz
        with open(SURROGATE_PMC_SUMMARY) as f:
            pmc = json.load(f)
        cycles = pmc.get("valu_issue_cycles_per_launch")
        if cycles:
            rate = cycles / (mlp_ms * 1e-3) / 1e9
            peak = 1024 * 2.4
            valu = {"achieved": round(rate, 1), "peak": peak, "unit": "G SIMD issue-cycles/s",
                    "frac": round(rate / peak, 4), "instructions_per_launch": pmc["sq"].get("SQ_INSTS_VALU"),
                    "transcendental_per_launch": pmc["sq"].get("SQ_INSTS_VALU_TRANS_F32")}
    line = {"metric": METRIC, "value": rec["surrogate_pipeline_points_per_s"], "unit": "points/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": rec["surrogate_pipeline_ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (surrogate: bf16 mfma)",
            "data": "synthetic",
            "config": {"workload": rec["workload"], "strands": N, "cvs_per_strand": 24, "attribute_programs": 4,
                       "threshold": rec["threshold"], "sample_stride": rec["sample_stride"],
                       "parallelism": "single-gpu"},
            "surrogate": rec,
            "roofline": {"bound": "mfma", "kernel": "surrogate_kernel (root features + MLP)",
                         "achieved": round(flops, 3), "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": round(flops / 2500.0, 6), "traffic": None,
                         "flop_per_root": SURROGATE_FLOP_PER_ROOT, "kernel_ms": mlp_ms, "valu": valu},
            "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


def bench_surrogate_sweep(args):
    """Acceptance and points/s of the surrogate pipeline against the exact
    one over relative thresholds 1e-3 .. 3e-2 and the programs' noise
    frequency scale 0.03 .. 1 (one trained network per scale); one JSON line
    per (scale, threshold)."""
    from nanoxgen_amd import workloads
    for fs in (0.03, 0.1, 0.3, 1.0):
        w = workloads.config4(frequency_scale=fs)
        for thr in (1e-3, 3e-3, 1e-2, 3e-2):
            rec = surrogate_record(args, w=w, threshold=thr, stride=64, steps=5, quiet=True)
            keep = ("frequency_scale", "threshold", "speedup_vs_exact", "exact_pipeline_points_per_s",
                    "surrogate_pipeline_points_per_s", "tiles_served", "tiles", "served_roots_within_threshold",
                    "served_roots_max_rel_error", "network_roots_within_threshold", "network_rel_error_p50_p99",
                    "train_loss_first_last")
            print(json.dumps({k: rec[k] for k in keep}), flush=True)
    return 0


def bench_batch(args):
    """Batched uneven form: nine descriptions (2,456,139 strands, 4..64 CVs,
    meshes of 2,048..199,712 triangles, per-description expression programs)
    in one captured graph: attribute programs, one guide preparation over
    all descriptions and the persistent strand kernel.  The tuning comes
    from the library's autotuner, run before the timed region."""
    import numpy as np
    import torch
    import nanoxgen_amd as nx
    from nanoxgen_amd import batch, expr, workloads

    t = time.perf_counter()
    ws = workloads.config5()
    blobs = [nx.build_asset(w.build) for w in ws]
    log(f"built {len(ws)} descriptions ({sum(b.nbytes for b in blobs)} B of assets) in "
        f"{time.perf_counter() - t:.1f} s")
    assets = [nx.DeviceAsset(b) for b in blobs]
    stream = torch.cuda.current_stream()
    full = batch.Batch()
    kern = batch.Batch()  # the same generation from the evaluated attributes (no programs)
    outs = []
    for w, a in zip(ws, assets):
        out = nx.PackedOutputs(w.params.strand_count, w.params.cvs_per_strand, roots=False, root_uvs=False)
        outs.append(out)
        k = full.add(a, w.params, out, octaves=w.octaves, exprs=expr.ExprSet(w.info["programs"]),
                     maps=expr.TexelMaps(w.info["maps"]))
        kern.add(a, w.params, out, octaves=w.octaves, attributes=full.entries[k]["attributes"],
                 workspace=full.entries[k]["workspace"])
    best, table = full.autotune(repeats=3)
    torch.cuda.synchronize()
    for tun, ms in sorted(table, key=lambda x: (x[1] is None, x[1] or 0.0)):
        log(f"  tuning {tun.as_tuple()}: {ms if ms is None else round(ms, 3)} ms")
    log(f"autotuner picked {best}")
    graph = full.graph(best)
    for _ in range(args.warmup):
        graph.launch(stream)
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        events[i][0].record(stream)
        graph.launch(stream)
        events[i][1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    step_ms = sorted(e[0].elapsed_time(e[1]) for e in events)
    # strand kernel (+ the guide preparation launch before it) from the
    # evaluated attributes, timed on the launch stream
    kev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(max(5, args.steps // 2))]
    for e in kev:
        e[0].record(stream)
        kern.generate(best, stream)
        e[1].record(stream)
    torch.cuda.synchronize()
    kern_ms = sorted(e[0].elapsed_time(e[1]) for e in kev)
    kernel_ms = sum(kern_ms) / len(kern_ms)
    points = full.points
    strands = full.strands
    alg_bytes = sum(w.params.strand_count * (w.params.cvs_per_strand * 16 + 4 + 16) for w in ws) + \
        sum(b.nbytes for b in blobs)
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "batch_kernel<%d, %d> (+ batch_prep_kernel)" % (best.waves_per_block, best.chunk_cvs),
                "kernel_ms_mean": round(kernel_ms, 4), "kernel_ms_median": round(kern_ms[len(kern_ms) // 2], 4),
                "step_ms_median": round(step_ms[len(step_ms) // 2], 4),
                "algorithmic_bytes_per_launch": alg_bytes}
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        oracle.lib()
        threads = baseline_threads()
        frac = 0.002
        cpu_pts, cpu_s, tries = 0, 0.0, 0
        while tries < 4:
            cpu_pts, cpu_s = 0, 0.0
            for w, b in zip(ws, blobs):
                s = np.arange(max(64, int(w.params.strand_count * frac)), dtype=np.uint32)
                t = time.perf_counter()
                attrs, _ = oracle.root_attributes(b, w.params, w.info["programs"], w.info["maps"], strands=s,
                                                  workers=threads)
                oracle.generate(b, w.params, octaves=w.octaves, strands=s, workers=threads, attributes=attrs)
                cpu_s += time.perf_counter() - t
                cpu_pts += len(s) * w.params.cvs_per_strand
            if cpu_s >= 0.5 * args.cpu_seconds:
                break
            frac = min(1.0, frac * args.cpu_seconds / max(cpu_s, 1e-3))
            tries += 1
        cpu = dict({"value": cpu_pts / cpu_s, "unit": "points/s", "cores": threads, "kind": "port",
                    "sample": f"the first {frac:.4f} of every description's strands ({cpu_pts} points), attribute "
                              f"programs + generation, {cpu_s:.2f} s; float restatement of the reference CPU path "
                              f"(oracle/), {threads} threads"}, **host_cpu())
    line = {"metric": METRIC, "value": points / (ms_per_step * 1e-3), "unit": "points/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"batch-{len(ws)}-descriptions-{strands}-strands-{points}-points",
                       "descriptions": [w.name for w in ws], "strands": strands, "points": points,
                       "tuning": dict(zip(["waves_per_block", "chunk_cvs", "tiles_per_claim", "blocks_per_cu",
                                           "order"], best.as_tuple())),
                       "launch": "one graph: attribute programs, batched guide preparation, persistent strand kernel",
                       "parallelism": "single-gpu"},
            "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    return 0


def train_surrogate(asset, w, es, maps, pairs, epochs):
    """A network trained on `pairs` exact-runtime attribute rows of roots drawn
    with a seed disjoint from the workload's (seed + 1000).  Returns (net,
    seconds, first and last loss)."""
    import torch
    from nanoxgen_amd import expr, surrogate
    t0 = time.perf_counter()
    train_p = w.params.copy(seed=w.params.seed + 1000, strand_count=pairs)
    feats = surrogate.root_features(asset, train_p, maps)
    targets = torch.empty((pairs, 4), device="cuda")
    expr.evaluate_root_attributes(asset, train_p, es, maps, targets)
    net = surrogate.Surrogate(seed=1)
    net.fit_norm(feats, targets)
    hist = net.train(feats, targets, epochs=epochs, batch=4096, lr=3e-3, log_every=256)
    torch.cuda.synchronize()
    return net, time.perf_counter() - t0, (hist[0][1], hist[-1][1])


def bench_batch_surrogate(args):
    """The batched uneven form with the learned surrogate on its descriptions
    of more than 20 ops (six of the nine).  The programs' noise3 argument
    scales are multiplied by 0.1 (the surrogate workload's scale; at 1.0 the
    networks serve no tile at 1e-3).  Each such description gets its own
    network, trained outside the timed region on 2^20 seed-disjoint exact
    pairs; in the batch its attributes are the network's, checked on one
    root per 64-root tile at the threshold, a tile that misses re-evaluated
    exactly.  The surrogate batch's graph is timed against the exact batch's
    graph of the same descriptions (same tuning, from the autotuner run on
    the exact batch); per description: tiles served, errors of the served
    roots against the exact attributes, and point counts."""
    import torch
    import nanoxgen_amd as nx
    from nanoxgen_amd import batch, expr, workloads

    ws = workloads.config5(frequency_scale=workloads.SURROGATE_FREQUENCY_SCALE)
    blobs = [nx.build_asset(w.build) for w in ws]
    assets = [nx.DeviceAsset(b) for b in blobs]
    stream = torch.cuda.current_stream()
    exact_b, sur_b = batch.Batch(), batch.Batch()
    outs_e, outs_s, trained = [], [], []
    for w, a in zip(ws, assets):
        es, maps = expr.ExprSet(w.info["programs"]), expr.TexelMaps(w.info["maps"])
        net = None
        if w.info["ops"] > 20:
            net, secs, loss = train_surrogate(a, w, es, maps, 1 << 20, args.train_epochs)
            trained.append((w.name, round(secs, 2), loss))
            log(f"{w.name}: trained in {secs:.1f} s, loss {loss[0]:.4f} -> {loss[1]:.2e}")
        n, c = w.params.strand_count, w.params.cvs_per_strand
        outs_e.append(nx.PackedOutputs(n, c, roots=False, root_uvs=False))
        outs_s.append(nx.PackedOutputs(n, c, roots=False, root_uvs=False))
        exact_b.add(a, w.params, outs_e[-1], octaves=w.octaves, exprs=es, maps=maps)
        sur_b.add(a, w.params, outs_s[-1], octaves=w.octaves, exprs=es, maps=maps, surrogate=net,
                  surrogate_threshold=args.threshold, surrogate_stride=args.sample_stride)
    best, _ = exact_b.autotune(repeats=3)
    torch.cuda.synchronize()
    log(f"autotuner picked {best}")
    g_exact, g_sur = exact_b.graph(best), sur_b.graph(best)

    def timed(g):
        for _ in range(args.warmup):
            g.launch(stream)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            g.launch(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / args.steps

    exact_ms = timed(g_exact)
    sur_ms = timed(g_sur)
    torch.cuda.synchronize()
    items = []
    for k, w in enumerate(ws):
        e_s, e_e = sur_b.entries[k], exact_b.entries[k]
        rec = {"description": w.name, "ops": w.info["ops"], "surrogate": e_s["surrogate"] is not None,
               "point_counts_exact": bool(torch.equal(outs_s[k].point_counts, outs_e[k].point_counts))}
        if e_s["surrogate"] is not None:
            redone, checked = sur_b.stats(k)
            tiles = e_s["surrogate_flags"].numel()
            exact = e_e["attributes"]
            rel = ((e_s["attributes"] - exact).abs() / exact.abs().clamp_min(1e-6)).max(1).values
            served = torch.repeat_interleave(e_s["surrogate_flags"] == 0, 64)[:w.params.strand_count]
            rs = rel[served]
            rec.update(tiles=tiles, tiles_served=tiles - redone, samples_checked=checked,
                       served_roots_within_threshold=float((rs <= args.threshold).float().mean()) if rs.numel()
                       else None,
                       served_roots_max_rel_error=float(rs.max()) if rs.numel() else None)
        items.append(rec)
        log(f"  {rec}")
    g_exact.close()
    g_sur.close()
    points = sur_b.points
    line = {"metric": METRIC, "value": points / (sur_ms * 1e-3), "unit": "points/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sur_ms, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 (surrogate: bf16 mfma)", "data": "synthetic",
            "config": {"workload": f"batch-surrogate-{len(ws)}-descriptions-{sur_b.strands}-strands-{points}-points"
                                   f"-f{workloads.SURROGATE_FREQUENCY_SCALE:g}",
                       "threshold": args.threshold, "sample_stride": args.sample_stride,
                       "tuning": dict(zip(["waves_per_block", "chunk_cvs", "tiles_per_claim", "blocks_per_cu",
                                           "order"], best.as_tuple())),
                       "parallelism": "single-gpu"},
            "exact_batch_ms": round(exact_ms, 4), "exact_batch_points_per_s": points / (exact_ms * 1e-3),
            "speedup_vs_exact": round(exact_ms / sur_ms, 4),
           "pipelined_steps": {"exact_pipeline_ms": round(pipe_ms["exact"], 4),
                               "surrogate_pipeline_ms": round(pipe_ms["surrogate"], 4),
                               "speedup_vs_exact": round(pipe_ms["exact"] / pipe_ms["surrogate"], 4)}, "items": items,
            "training": [{"description": n, "seconds": s, "loss_first_last": l} for n, s, l in trained],
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


CUT_PMC_SUMMARY = os.path.join(ROOT, "profiles", "cut_search_kernel_plane2M_pmc.json")
RUNTIME_PMC_SUMMARY = os.path.join(ROOT, "profiles", "runtime_chain_plane2M_classic_expr_pmc.json")
PLAN_PMC_SUMMARY = os.path.join(ROOT, "profiles", "runtime_chain_plane2M_classic_plan_pmc.json")


def bench_cdf(args):
    """The plane workload (2,000,000 strands x 16 CVs, two octaves) with the
    opt-in inverse-CDF root sampler.  The sampler is built once from the
    asset (timed apart); one step = nxg_generate_packed_cdf: the CDF roots
    kernel (two-level search, roots written to the output), guide prep, the
    strand order by the sampled triangles, the strand kernel.  The alias
    mode with the same outputs (roots on) is timed in the same process."""
    import numpy as np
    import torch
    import nanoxgen_amd as nx
    from nanoxgen_amd import workloads

    w = workloads.config2()
    N, Cv = w.params.strand_count, w.params.cvs_per_strand
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    stream = torch.cuda.current_stream()
    nx.CdfSampler(asset, stream=stream)  # first build: allocation and module load
    build_ms = []
    for _ in range(5):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e[0].record(stream)
        sampler = nx.CdfSampler(asset, stream=stream)
        e[1].record(stream)
        torch.cuda.synchronize()
        build_ms.append(e[0].elapsed_time(e[1]))
    out = nx.PackedOutputs(N, Cv, roots=True, root_uvs=False, point_counts=True)
    ws = torch.empty(asset.workspace_bytes(w.params), dtype=torch.uint8, device="cuda")

    def run(use_cdf):
        def step():
            nx.generate_packed(asset, w.params, out, octaves=w.octaves, workspace=ws, stream=stream,
                               sampler=sampler if use_cdf else None)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
        t0 = time.perf_counter()
        for e in ev:
            e[0].record(stream)
            step()
            e[1].record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / args.steps
        ms = sorted(e[0].elapsed_time(e[1]) for e in ev)
        return wall, sum(ms) / len(ms), ms[len(ms) // 2]

    alias_wall, alias_mean, _ = run(False)
    wall, call_mean, call_median = run(True)
    points = N * Cv
    T = asset.header.triangle_count
    # outputs + counts + roots written by the sampler and read back by the
    # strand kernel + the blob + the prefix sums read once
    alg = points * 16 + N * 4 + 2 * N * 48 + blob.nbytes + T * 8
    achieved = alg / (call_mean * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "the whole call: cdf_roots_kernel + guide_prep_kernel + strand order + strand_kernel<true, 2>",
                "kernel_ms_mean": round(call_mean, 4), "kernel_ms_median": round(call_median, 4),
                "algorithmic_bytes_per_launch": alg}
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        pre = oracle.cdf_prefix(blob)
        threads = baseline_threads()

        def cpu_run(s):
            t = time.perf_counter()
            roots = oracle.cdf_roots(blob, w.params.seed, s, pre)
            oracle.generate(blob, w.params.copy(strand_count=s), octaves=w.octaves, roots=roots, workers=threads)
            return time.perf_counter() - t

        s, dt, runs = sampled_seconds(cpu_run, N, args.cpu_seconds)
        cpu = dict({"value": s * Cv / dt, "unit": "points/s", "cores": threads, "kind": "port",
                    "sample": f"strands 0..{s - 1} ({s * Cv} points): the restatement's CDF roots "
                              f"(oracle/oracle_cdf.c, one thread) + generation from them ({threads} threads), "
                              f"median {dt:.3f} s of {runs} run(s)"}, **host_cpu())
    line = {"metric": METRIC, "value": points / (wall * 1e-3), "unit": "points/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "plane-cdf-sampler", "input": w.name, "strands": N, "cvs_per_strand": Cv,
                       "triangles": T, "sampler": "inverse CDF (opt-in)", "parallelism": "single-gpu"},
            "sampler_build_ms_median": round(sorted(build_ms)[len(build_ms) // 2], 4),
            "alias_same_outputs": {"ms_per_step": round(alias_wall, 4), "call_ms_mean": round(alias_mean, 4),
                                   "points_per_s": points / (alias_wall * 1e-3)},
            "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    return 0


def chain_counters(path, chain_ms):
    """(HBM traffic per chain, VALU roofline dict) from a committed rocprofv3
    counter summary of the same chain, or (None, None)."""
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        pmc = json.load(f)
    valu = None
    insts = pmc.get("valu_instructions_per_chain")
    if insts:
        valu = valu_issue(insts, chain_ms, "instructions_per_chain")
    return pmc.get("hbm_bytes_per_chain"), valu


def bench_classic(args):
    """Classic CutFX (nxg_classic_cut_apply) over the plane workload's
    generated curves: 2,000,000 strands x 16 CVs with point counts, roots and
    root uvs, amount = 0.25 x cLength.  One step = the whole chain (search,
    two scan launches, rebuild + compaction); the curves are generated once,
    before the timed region, and stay resident in HBM."""
    import numpy as np
    import torch
    import nanoxgen_amd as nx
    from nanoxgen_amd import workloads
    from nanoxgen_amd.classic import ClassicCutOutputs, classic_cut, workspace_bytes

    scale, offset = 0.25, 0.0
    w = workloads.config2()
    N, Cv = w.params.strand_count, w.params.cvs_per_strand
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    gen = nx.PackedOutputs(N, Cv)
    nx.generate_packed(asset, w.params, gen, octaves=w.octaves)
    out = ClassicCutOutputs(N, Cv)
    ws = torch.empty(workspace_bytes(N), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    log(f"generated {N} x {Cv} curves")

    def step():
        classic_cut(gen.points, Cv, scale, offset, point_counts=gen.point_counts, roots=gen.roots,
                    root_uvs=gen.root_uvs, outputs=out, workspace=ws, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        events[i][0].record(stream)
        step()
        events[i][1].record(stream)
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    chain_ms = sorted(e[0].elapsed_time(e[1]) for e in events)
    chain_mean = sum(chain_ms) / len(chain_ms)
    kept = out.kept()
    assert out.status() == 0
    # algorithmic bytes of one chain: every input strand read (points, count,
    # root, uv), every kept strand written, 4 B of cut parameter written and
    # read per strand, the tile counts negligible
    per_in, per_out = Cv * 16 + 4 + 48 + 8, Cv * 16 + 4 + 48 + 8
    alg = N * per_in + kept * per_out + N * 8
    achieved = alg / (chain_mean * 1e-3) / 1e9
    traffic, valu = chain_counters(CUT_PMC_SUMMARY, chain_mean)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": "cut chain: cut_table + cut_search_kernel<true> + cut_scan_tiles + cut_scan_groups + "
                          "cut_write_held_kernel<16> (one C-ABI call)",
                "kernel_ms_mean": round(chain_mean, 4), "kernel_ms_median": round(chain_ms[len(chain_ms) // 2], 4),
                "algorithmic_bytes_per_launch": alg, "valu": valu}
    # the renderer endpoint pass on the cut's output, its strand count read on the device
    from nanoxgen_amd.classic import add_endpoints
    ends = torch.empty((N * (Cv + 2), 4), dtype=torch.float32, device="cuda")
    end_counts = torch.empty(N, dtype=torch.int32, device="cuda")
    for _ in range(3):
        add_endpoints(out.points, Cv, strands=N, device_strand_count=out.words, out=ends, point_counts=end_counts,
                      stream=stream)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    for e in ev:
        e[0].record(stream)
        add_endpoints(out.points, Cv, strands=N, device_strand_count=out.words, out=ends, point_counts=end_counts,
                      stream=stream)
        e[1].record(stream)
    torch.cuda.synchronize()
    end_ms = sorted(e[0].elapsed_time(e[1]) for e in ev)
    end_mean = sum(end_ms) / len(end_ms)
    end_bytes = kept * (Cv * 16 + (Cv + 2) * 16 + 4)
    roofline["endpoints"] = {"kernel": "endpoints_kernel", "ms_mean": round(end_mean, 4),
                             "algorithmic_bytes": end_bytes,
                             "achieved_gbs": round(end_bytes / (end_mean * 1e-3) / 1e9, 1),
                             "frac": round(end_bytes / (end_mean * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = classic_cpu_baseline(gen.to_host(), Cv, scale, offset, args.cpu_seconds, baseline_threads())
    line = {"metric": METRIC, "value": kept * Cv / (ms_per_step * 1e-3), "unit": "points/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "classic-cut-plane2M", "input": w.name, "strands": N, "cvs_per_strand": Cv,
                       "kept_strands": kept, "amount": f"{scale} * cLength + {offset}",
                       "parallelism": "single-gpu"},
            "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    return 0


# The Classic plan of the classic_expr workload: authored form (the
# reference's .xgen attributes) and bound form (map -> input columns).
CLASSIC_EXPR_AUTHORED = dict(length="1 + 0.5 * map('len.ptx')", width="0.004 + 0.01 * map('w.ptx')",
                             taper="0.6", taper_start="0.25", width_ramp="rampUI(0,1,1:0.4,0.7,2:1,0.2,3)",
                             cut_amount="$cLength * (0.2 + 0.5 * rand(3)) + $cWidth")
CLASSIC_EXPR_BOUND = dict(CLASSIC_EXPR_AUTHORED, length="1 + 0.5 * $__nxg_map_0",
                          width="0.004 + 0.01 * $__nxg_map_1")
CLASSIC_EXPR_ATTRS = (("length", "length"), ("width", "width"), ("taper", "taper"), ("taper_start", "taperStart"),
                      ("width_ramp", "widthRamp"))


# The classic_plan workload's ClumpingFX module (toward the plane's guides).
CLASSIC_PLAN_CLUMP = dict(mask="1", clump="0.6 + 0.3 * rand()", clumpScale="rampUI(0,0,1:1,0.2,1)", noise="0.3",
                          noiseScale="1", noiseFrequency="2", noiseCorrelation="20")


# The classic_noise workload's NoiseFX module (after the clump, before the cut).
CLASSIC_NOISE = dict(mask="1", magnitude="0.05 + 0.05 * rand()", magnitudeScale="rampUI(0,0,1:1,1,1)",
                     frequency="3", correlation="20", preserveLength="50")


def classic_expr_xgen(with_clump=False, with_noise=False):
    """The classic_expr plan (and, with_clump, the classic_plan clump before
    the cut) as a Classic .xgen description."""
    a = CLASSIC_EXPR_AUTHORED
    s = ("FileVersion 18\nPalette\n\tname\tpalette\n\tendAttrs\nDescription\n\tname\tplane\n"
         "\tdescriptionId\t2\n\tendAttrs\nSplinePrimitive\n")
    s += "".join(f"\t{attr}\t{a[key]}\n" for key, attr in CLASSIC_EXPR_ATTRS) + "\tendAttrs\n"
    if with_clump:
        s += "ClumpingFXModule\n\tactive\ttrue\n\tname\tclump\n"
        s += "".join(f"\t{k}\t{v}\n" for k, v in CLASSIC_PLAN_CLUMP.items()) + "\tendAttrs\n"
    if with_noise:
        s += "NoiseFXModule\n\tactive\ttrue\n\tname\tnoise\n\tmode\t0\n"
        s += "".join(f"\t{k}\t{v}\n" for k, v in CLASSIC_NOISE.items()) + "\tendAttrs\n"
    s += f"CutFXModule\n\tactive\ttrue\n\tname\tcut\n\tamount\t{a['cut_amount']}\n\trebuildType\t1\n\tendAttrs\n"
    return s + "\tActive\tSplinePrimitive\n"


def bench_classic_expr(args, with_clump=False, with_noise=False):
    """The Classic runtime with primitive expressions
    (nxg_classic_runtime_apply) over the plane workload's generated curves:
    2,000,000 strands x 16 CVs, a length and a width expression over two
    input columns, taper 0.6 from 0.25, a three-key widthRamp, and a CutFX
    whose amount calls rand() through per-strand SeExpr prefixes; with_clump
    (classic_plan) adds a ClumpingFX module with expressions before the cut,
    each strand clumped to its stencil's slot-0 guide among the
    description's 4,096 guides.  One step = the whole chain (sample tables,
    uniform programs, guide pass, search with every expression, two scan
    launches, write with the width profile + compaction); inputs are
    resident in HBM before the timed region."""
    import numpy as np
    import torch
    import nanoxgen_amd as nx
    from nanoxgen_amd import abi, workloads
    from nanoxgen_amd.classic import ClassicCutOutputs, ClassicPlan, classic_runtime, runtime_workspace_bytes

    w = workloads.config2()
    N, Cv = w.params.strand_count, w.params.cvs_per_strand
    blob = nx.build_asset(w.build)
    gen = nx.PackedOutputs(N, Cv, stencil=with_clump)
    nx.generate_packed(nx.DeviceAsset(blob), w.params, gen, octaves=w.octaves)
    rng = np.random.default_rng(17)
    columns = rng.uniform(0.0, 1.0, (N, 2)).astype(np.float32)
    prefixes = rng.integers(0, 2 ** 32, N, dtype=np.uint64).astype(np.uint32)
    plan = ClassicPlan(**CLASSIC_EXPR_BOUND, clump=CLASSIC_PLAN_CLUMP if with_clump else None,
                       noise=CLASSIC_NOISE if with_noise else None, description_id=2,
                       description_name="plane", ptex_count=2)
    cols_d = torch.from_numpy(columns).cuda()
    pre_d = torch.from_numpy(prefixes.view(np.int32)).cuda()
    clump_kw, guide_host, guide_count = {}, None, 0
    if with_clump:
        h = abi.header_of(blob)
        guide_count = h.guide_count
        recs = blob[h.guides_offset:h.guides_offset + 56 * guide_count].view(abi.GUIDE_DTYPE)
        cvs3 = blob[h.guide_cvs_offset:h.guide_cvs_offset + 12 * h.guide_cv_count].view(np.float32).reshape(-1, 3)
        guide_host = dict(axes=cvs3.copy(), normals=recs["root_normal"].astype(np.float32),
                          tangents=np.zeros((guide_count, 3), np.float32))
        d = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
        guide_uvs = rng.uniform(0, 1, (guide_count, 2)).astype(np.float32)
        guide_prefixes = rng.integers(0, 2 ** 32, guide_count, dtype=np.uint64).astype(np.uint32)
        clump_kw = dict(strand_guides=gen.stencil_indices[:, 0].contiguous(),
                        guides={k: d(v) for k, v in guide_host.items()}, guide_uvs=d(guide_uvs),
                        guide_random_prefixes=torch.from_numpy(guide_prefixes.view(np.int32)).cuda())
        guide_host.update(uvs=guide_uvs, prefixes=guide_prefixes)
    out = ClassicCutOutputs(N, Cv)
    ws = torch.empty(runtime_workspace_bytes(N, Cv, guide_count, plan), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    log(f"generated {N} x {Cv} curves; plan image {plan.info.byte_size} B; {guide_count} clump guides")

    tangents = np.tile(np.float32([1.0, 0.0, 0.0]), (N, 1))  # per-strand surface tangents of the plane
    if with_noise:
        clump_kw["surface_tangents"] = torch.from_numpy(tangents).cuda()

    def step():
        classic_runtime(gen.points, Cv, plan, roots=gen.roots, root_uvs=gen.root_uvs, point_counts=gen.point_counts,
                        random_prefixes=pre_d, columns=cols_d, outputs=out, workspace=ws, stream=stream, **clump_kw)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        events[i][0].record(stream)
        step()
        events[i][1].record(stream)
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    chain_ms = sorted(e[0].elapsed_time(e[1]) for e in events)
    chain_mean = sum(chain_ms) / len(chain_ms)
    kept = out.kept()
    assert out.status() == 0
    # algorithmic bytes of one chain: every input strand's points, count,
    # root, uv, two input columns and prefix read once; the 4 B cut parameter
    # and 32 B state written by the search and read by the write; every kept
    # strand's points, count, root and uv written
    per_in = Cv * 16 + 4 + 48 + 8 + 8 + 4
    per_out = Cv * 16 + 4 + 48 + 8
    alg = N * per_in + N * 36 * 2 + kept * per_out
    if with_clump:  # the strand's guide index, its guide's axes, the clumped copy written and read back
        alg += N * 4 + guide_count * Cv * (12 + 32) + N * Cv * 16 * 2
    achieved = alg / (chain_mean * 1e-3) / 1e9
    # (the counter summaries cover the variant without the NoiseFX step)
    traffic, valu = (None, None) if with_noise else chain_counters(
        PLAN_PMC_SUMMARY if with_clump else RUNTIME_PMC_SUMMARY, chain_mean)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": "runtime chain: cut_table + classic_uniform_kernel + "
                          + ("runtime_guide_kernel + " if with_clump else "") + f"runtime_search_kernel<true, {'true' if with_noise else 'false'}> + "
                          "cut_scan_tiles + cut_scan_groups + runtime_write_held_kernel<16> (one C-ABI call)",
                "kernel_ms_mean": round(chain_mean, 4), "kernel_ms_median": round(chain_ms[len(chain_ms) // 2], 4),
                "algorithmic_bytes_per_launch": alg, "valu": valu}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = classic_expr_cpu_baseline(gen.to_host(), Cv, columns, prefixes, args.cpu_seconds, baseline_threads(),
                                        guide_host, tangents if with_noise else None)
    plan_text = dict(CLASSIC_EXPR_AUTHORED, **({"clump": CLASSIC_PLAN_CLUMP} if with_clump else {}),
                     **({"noise": CLASSIC_NOISE} if with_noise else {}))
    line = {"metric": METRIC, "value": kept * Cv / (ms_per_step * 1e-3), "unit": "points/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "classic-noise-plane2M" if with_noise else
                       "classic-plan-plane2M" if with_clump else "classic-expressions-plane2M",
                       "input": w.name, "strands": N, "cvs_per_strand": Cv, "kept_strands": kept,
                       "clump_guides": guide_count, "plan": plan_text, "parallelism": "single-gpu"},
            "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    return 0


def classic_expr_cpu_baseline(host, Cv, columns, prefixes, target_s, threads, guides=None, tangents=None,
                              use_ref=False):
    """The oracle's restatement of the Classic runtime chain
    (oracle/oracle_classic.c, strands over `threads` workers, then the stable
    compaction) on the first S strands of `host` (the generated curves as
    host arrays) with the same plan; use_ref times the reference's parser,
    plan compiler and apply_xgen_classic_float_runtime_plan_cpu (oracle/_ref,
    its own thread pool) instead.  guides (host arrays) adds the clump."""
    import oracle
    from nanoxgen_amd.classic import ClassicPlan
    pts, counts, roots, uvs = host["points"], host["point_counts"], host["roots"], host["root_uvs"]
    N = len(counts)
    text = classic_expr_xgen(guides is not None, tangents is not None)
    plan = ClassicPlan(**CLASSIC_EXPR_BOUND, clump=CLASSIC_PLAN_CLUMP if guides is not None else None,
                       noise=CLASSIC_NOISE if tangents is not None else None,
                       description_id=2, description_name="plane", ptex_count=2)

    def clump_kw(s):
        kw = {} if tangents is None else dict(surface_tangents=tangents[:s])
        if guides is None:
            return kw
        return dict(kw, strand_guides=host["stencil_indices"][:s, 0], guide_uvs=guides["uvs"],
                    guide_random_prefixes=guides["prefixes"],
                    guides=dict(axes=guides["axes"], normals=guides["normals"], tangents=guides["tangents"]))

    def run(s):
        if use_ref:
            r = oracle.ref_classic_runtime(text, pts[:s * Cv], Cv, roots=roots[:s], counts=counts[:s], uvs=uvs[:s],
                                           random_prefixes=prefixes[:s], columns=columns[:s], **clump_kw(s))
            return r["seconds"]
        t = time.perf_counter()
        r = oracle.classic_runtime(pts[:s * Cv], Cv, plan, roots=roots[:s], counts=counts[:s], uvs=uvs[:s],
                                   random_prefixes=prefixes[:s], columns=columns[:s], workers=threads,
                                   **clump_kw(s))
        dt = time.perf_counter() - t
        assert r["strand_count"] is not None, r["error_bits"]
        return dt

    s, dt, runs = sampled_seconds(run, N, target_s)
    return baseline_record(s, Cv, dt, runs, threads, use_ref, "the same plan",
                           "apply_xgen_classic_float_runtime_plan_cpu")


def baseline_record(s, Cv, dt, runs, threads, use_ref, what, ref_entry):
    """A Classic workload's cpu_baseline object: the restatement ("port", on
    `threads` workers) or the reference build ("reference", its own pool)."""
    who = (f"the reference's {ref_entry} (oracle/_ref, its own thread pool, seconds measured inside the call)"
           if use_ref else f"the oracle's restatement (oracle/oracle_classic.c, oracle_clump.c), {threads} threads; "
                           "oracle/standin_check.py times it against the reference build on the same host")
    return dict({"value": s * Cv / dt, "unit": "points/s",
                 "cores": len(os.sched_getaffinity(0)) if use_ref else threads,
                 "kind": "reference" if use_ref else "port",
                 "sample": f"strands 0..{s - 1} of the generated curves ({s * Cv} points), {what}, "
                           f"median {dt:.3f} s of {runs} run(s); {who}"}, **host_cpu())


CLUMP_PARAMS = dict(mask=0.8, amount=0.7, clump_scale=0.1, noise=0.3, noise_scale=1.0, noise_frequency=2.0,
                    noise_correlation=20.0)


def bench_clump(args):
    """ClumpingFX (nxg_classic_clump_apply) over the plane workload's
    generated curves: 2,000,000 strands x 16 CVs, each clumped to the nearest
    guide of its stencil (slot 0) among the description's 4,096 guides of 16
    CVs, constant module expressions with clump noise on.  One step = the
    guide pass (render axes, lengths, noise axes) + the strand pass."""
    import numpy as np
    import torch
    import nanoxgen_amd as nx
    from nanoxgen_amd import abi, workloads
    from nanoxgen_amd.classic import clump, clump_workspace_bytes

    w = workloads.config2()
    N, Cv = w.params.strand_count, w.params.cvs_per_strand
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    gen = nx.PackedOutputs(N, Cv, stencil=True)
    nx.generate_packed(asset, w.params, gen, octaves=w.octaves)
    h = abi.header_of(blob)
    recs = blob[h.guides_offset:h.guides_offset + 56 * h.guide_count].view(abi.GUIDE_DTYPE)
    assert np.all(recs["cv_count"] == Cv) and np.array_equal(recs["first_cv"], np.arange(h.guide_count) * Cv)
    cvs3 = blob[h.guide_cvs_offset:h.guide_cvs_offset + 12 * h.guide_cv_count].view(np.float32).reshape(-1, 3)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    guides = dict(axes=d(cvs3), normals=d(recs["root_normal"]), tangents=d(np.zeros((h.guide_count, 3))))
    strand_guides = gen.stencil_indices[:, 0].contiguous()
    out = torch.empty_like(gen.points)
    ws = torch.empty(clump_workspace_bytes(h.guide_count, Cv), dtype=torch.uint8, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    log(f"generated {N} x {Cv} curves; {h.guide_count} clump guides")

    def step():
        clump(gen.points, Cv, strand_guides, guides, CLUMP_PARAMS, out=out, status=status, workspace=ws,
              stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        events[i][0].record(stream)
        step()
        events[i][1].record(stream)
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    call_ms = sorted(e[0].elapsed_time(e[1]) for e in events)
    call_mean = sum(call_ms) / len(call_ms)
    assert int(status.item()) == 0
    # algorithmic bytes: every strand read and written, its guide index read,
    # the guides' CVs, normals and tangents read once
    alg = N * Cv * 32 + N * 4 + h.guide_count * (Cv * 12 + 24)
    achieved = alg / (call_mean * 1e-3) / 1e9
    traffic = valu = None
    pmc_path = os.path.join(ROOT, "profiles", "clump_strand_kernel_plane2M_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            pmc = json.load(f)
        traffic = pmc.get("hbm_bytes_per_call")
        insts = pmc.get("valu_instructions_per_call")
        if insts:
            valu = valu_issue(insts, call_mean, "instructions_per_call")
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": "clump call: cut_table_kernel + clump_guide_kernel + clump_strand_kernel<16> "
                          "(one C-ABI call)",
                "kernel_ms_mean": round(call_mean, 4), "kernel_ms_median": round(call_ms[len(call_ms) // 2], 4),
                "algorithmic_bytes_per_launch": alg, "valu": valu}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = clump_cpu_baseline(gen.to_host(), Cv, cvs3, recs, args.cpu_seconds, baseline_threads())
    line = {"metric": METRIC, "value": N * Cv / (ms_per_step * 1e-3), "unit": "points/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "classic-clump-plane2M", "input": w.name, "strands": N, "cvs_per_strand": Cv,
                       "clump_guides": int(h.guide_count), "params": CLUMP_PARAMS, "parallelism": "single-gpu"},
            "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    return 0


def clump_cpu_baseline(host, Cv, cvs3, recs, target_s, threads, use_ref=False):
    """The oracle's restatement of ClumpingFX with constant module
    expressions (oracle/oracle_clump.c, strands over `threads` workers) on
    the first S strands of `host`, each clumped to its stencil's slot-0
    guide; use_ref times the reference's float runtime with the same module
    (oracle/_ref, its own thread pool; its time includes the shim's copies)."""
    import numpy as np
    import oracle
    pts = host["points"]
    N = len(host["point_counts"])
    sg = host["stencil_indices"][:, 0].copy()
    guides = dict(axes=cvs3, normals=np.ascontiguousarray(recs["root_normal"], np.float32),
                  tangents=np.zeros((len(recs), 3), np.float32))
    gradients = oracle.seexpr_gradients()

    def run(s):
        t = time.perf_counter()
        if use_ref:
            oracle.ref_clump_apply(pts[:s * Cv], Cv, sg[:s], guides, CLUMP_PARAMS)
        else:
            r = oracle.clump_apply(pts[:s * Cv], Cv, sg[:s], guides, CLUMP_PARAMS, gradients, workers=threads)
            assert r is not None
        return time.perf_counter() - t

    s, dt, runs = sampled_seconds(run, N, target_s)
    return baseline_record(s, Cv, dt, runs, threads, use_ref, "one ClumpingFX module",
                           "float runtime with one ClumpingFX module")


def bench_transport(args):
    """Renderer transport over the plane workload's generated curves
    (2,000,000 strands x 16 CVs, as the reference's SoA GeneratedCurves:
    Vec3 points + per-point widths + root uvs): build_curve_batches into 64K
    batches with a transform and one motion sample (nxg_curve_payload_build,
    the timed step), then the .nxc curve cache of the packed curves assembled
    on the GPU (reported beside it: one download plus the serial host hash)."""
    import numpy as np
    import torch
    import nanoxgen_amd as nx
    from nanoxgen_amd import workloads
    from nanoxgen_amd.cache import build_curve_cache
    from nanoxgen_amd.payload import curve_payload

    w = workloads.config2()
    N, Cv = w.params.strand_count, w.params.cvs_per_strand
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    gen = nx.PackedOutputs(N, Cv)
    nx.generate_packed(asset, w.params, gen, octaves=w.octaves)
    pts = gen.points[:, :3].contiguous()
    widths = (gen.points[:, 3] * 2.0).contiguous()
    base = dict(fixed=Cv, points=pts, widths=widths, root_uvs=gen.root_uvs)
    motion = [dict(fixed=Cv, points=(pts + 0.01).contiguous(), time=0.5)]
    transform = (1.0, 0.0, 0.0, 5.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, -3.0)
    stream = torch.cuda.current_stream()

    # the output buffers and workspace are made once and reused by every step,
    # as a renderer reuses its frame buffers
    reuse = curve_payload(base, motion, transform=transform, stream=stream)

    def step():
        return curve_payload(base, motion, transform=transform, stream=stream, out=reuse)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e in evs:
        e[0].record(stream)
        out = step()
        e[1].record(stream)
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    call = sorted(e[0].elapsed_time(e[1]) for e in evs)
    call_mean = sum(call) / len(call)
    assert int(out["status"].item()) == 0
    P = N * Cv
    alg = P * (12 + 4 + 12) + N * 8 + P * (16 + 12) + N * (4 + 8)
    achieved = alg / (call_mean * 1e-3) / 1e9
    t1 = time.perf_counter()
    cache = build_curve_cache(gen.point_counts, gen.points, patch_uvs=gen.root_uvs)
    cache_ms = (time.perf_counter() - t1) * 1e3
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "payload call: offsets scans + batch_points_kernel + payload_points_kernel + "
                          "payload_motion_kernel (one C-ABI call)",
                "kernel_ms_mean": round(call_mean, 4), "kernel_ms_median": round(call[len(call) // 2], 4),
                "algorithmic_bytes_per_launch": alg, "valu": None,
                "curve_cache": {"bytes": int(cache.nbytes), "ms": round(cache_ms, 2),
                                "what": "GPU assembly + one D2H download + host content hash and validation"}}
    line = {"metric": METRIC, "value": P / (ms_per_step * 1e-3), "unit": "points/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "transport-payload-plane2M", "input": w.name, "strands": N, "cvs_per_strand": Cv,
                       "batches": len(out["batch_points"]) - 1, "motion_samples": 1, "parallelism": "single-gpu"},
            "roofline": roofline, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


def classic_cpu_baseline(host, Cv, scale, offset, target_s, threads, use_ref=False):
    """The oracle's restatement of the CutFX chain (oracle/oracle_classic.c,
    strands over `threads` workers, then the stable compaction) on the first
    S strands of `host`; use_ref times the reference's own
    apply_xgen_classic_float_runtime_plan_cpu (oracle/_ref, its own pool)."""
    import oracle
    pts, counts, roots, uvs = host["points"], host["point_counts"], host["roots"], host["root_uvs"]
    N = len(counts)

    def run(s):
        if use_ref:
            r = oracle.ref_classic_cut(pts[:s * Cv], Cv, scale, offset, counts=counts[:s], roots=roots[:s],
                                       uvs=uvs[:s])
            return r["seconds"]
        t = time.perf_counter()
        r = oracle.classic_cut(pts[:s * Cv], Cv, scale, offset, counts=counts[:s], roots=roots[:s], uvs=uvs[:s],
                               workers=threads)
        dt = time.perf_counter() - t
        assert r is not None
        return dt

    s, dt, runs = sampled_seconds(run, N, target_s)
    return baseline_record(s, Cv, dt, runs, threads, use_ref, f"one CutFX at {scale} * cLength",
                           "apply_xgen_classic_float_runtime_plan_cpu")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 timed steps of the plane take about 75 ms: enough that the first
    # step's launch latency (the queue starts empty) is amortised
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--workload", default="plane",
                    choices=["plane", "plane_cdf", "expr", "surrogate", "surrogate_sweep", "batch",
                             "batch_surrogate", "classic",
                             "classic_expr", "classic_plan", "classic_noise", "clump", "transport"],
                    help="plane: 2,000,000 strands x 16 CVs, two octaves (default); "
                         "plane_cdf: the same with the opt-in inverse-CDF root sampler; "
                         "expr: 1,000,000 icosphere roots x 24 CVs with three 40-op attribute programs; "
                         "surrogate: the plane with 1,000,000 roots x 24 CVs and four 40-op attribute programs, "
                         "the trained surrogate + sampled exact check against the exact runtime; "
                         "surrogate_sweep: that comparison over thresholds and program noise frequencies; "
                         "batch: nine uneven descriptions (2,456,139 strands, 4..64 CVs) in one captured graph; "
                         "batch_surrogate: the same at program noise frequency scale 0.1, the surrogate on the "
                         "descriptions of more than 20 ops, against the exact batch; "
                         "classic: Classic CutFX (0.25 x cLength) over the plane's 2,000,000 generated strands; "
                         "classic_expr: the same strands through a Classic plan of length, width, taper, "
                         "taperStart, widthRamp and CutFX-amount expressions; "
                         "classic_plan: the same plan with an expression ClumpingFX toward the guides before the cut; "
                         "clump: ClumpingFX toward the plane's 4,096 guides over the same strands; "
                         "transport: 64K renderer batches (transform, motion) and the .nxc cache of the same strands")
    ap.add_argument("--threshold", type=float, default=1e-3, help="surrogate: relative attribute error bound")
    ap.add_argument("--sample-stride", type=int, default=64, help="surrogate: exact check of every n-th root")
    ap.add_argument("--train-epochs", type=int, default=200, help="surrogate: epochs over the 2^20 training pairs")
    ap.add_argument("--no-surrogate", action="store_true", help="plane: leave out the surrogate record")
    ap.add_argument("--serial-steps", action="store_true",
                    help="plane: run the steps one after another on one stream (no overlapped preparation)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="instead of the benchmark, run a plain device copy of this many MiB "
                         "(calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE against known bytes)")
    args = ap.parse_args()
    if args.copy_calibration_mib:
        import torch
        n = args.copy_calibration_mib << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "launches": 3}), flush=True)
        return 0
    if args.workload == "plane_cdf" and args.gpus == 1:
        return bench_cdf(args)
    if args.workload == "surrogate" and args.gpus == 1:
        return bench_surrogate(args)
    if args.workload == "surrogate_sweep" and args.gpus == 1:
        return bench_surrogate_sweep(args)
    if args.workload == "batch" and args.gpus == 1:
        return bench_batch(args)
    if args.workload == "batch_surrogate" and args.gpus == 1:
        return bench_batch_surrogate(args)
    if args.workload == "classic" and args.gpus == 1:
        return bench_classic(args)
    if args.workload == "classic_expr" and args.gpus == 1:
        return bench_classic_expr(args)
    if args.workload == "classic_plan" and args.gpus == 1:
        return bench_classic_expr(args, with_clump=True)
    if args.workload == "classic_noise" and args.gpus == 1:
        return bench_classic_expr(args, with_clump=True, with_noise=True)
    if args.workload == "clump" and args.gpus == 1:
        return bench_clump(args)
    if args.workload == "transport" and args.gpus == 1:
        return bench_transport(args)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(strands are independent: a multi-GPU run shards strand ranges, one process per GPU)", file=sys.stderr)
        return 2

    import torch
    import nanoxgen_amd as nx
    from nanoxgen_amd import workloads

    w = workloads.config2() if args.workload == "plane" else workloads.config3()
    N, Cv = w.params.strand_count, w.params.cvs_per_strand
    t = time.perf_counter()
    blob = nx.build_asset(w.build)
    log(f"built {w.name}: {blob.nbytes} B in {time.perf_counter() - t:.2f} s")
    asset = nx.DeviceAsset(blob)
    out = nx.PackedOutputs(N, Cv, roots=False, root_uvs=False, point_counts=True)
    ws = torch.empty(asset.workspace_bytes(w.params), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    attrs = exprs = maps = None
    if args.workload == "expr":
        from nanoxgen_amd import expr
        exprs = expr.ExprSet(w.info["programs"])
        maps = expr.TexelMaps(w.info["maps"])
        attrs = torch.empty((N, 4), dtype=torch.float32, device="cuda")

    # plane: consecutive steps through GenerationPipeline, each step's
    # preparation (guide state, offset table, strand order) on a side stream
    # overlapping the previous step's strand kernel; every step runs every
    # kernel.  The serial schedule is timed beside it for the record.
    pipe = None
    if not args.serial_steps:
        from nanoxgen_amd.pipeline import GenerationPipeline
        ws = None
        attribute_fn = None
        if attrs is not None:  # expr: each step's attributes beside its preparation, on the side stream
            def attribute_fn(params, out_attrs, side):
                expr.evaluate_root_attributes(asset, params, exprs, maps, out_attrs, stream=side)
        pipe = GenerationPipeline(asset, w.params, octaves=w.octaves, stream=stream, attribute_fn=attribute_fn)

    def step(ev=None):
        if pipe is not None:
            pipe.submit(out, events=ev)
            return
        if ev:
            ev[0].record(stream)
        if attrs is not None:
            expr.evaluate_root_attributes(asset, w.params, exprs, maps, attrs, stream=stream)
        else:
            nx.prepare_packed(asset, w.params, ws, stream=stream)
        if ev:
            ev[1].record(stream)
            ev[2].record(stream)
        if attrs is not None:
            nx.generate_packed(asset, w.params, out, octaves=w.octaves, workspace=ws, stream=stream,
                               attributes=attrs)
        else:
            nx.generate_packed_prepared(asset, w.params, out, ws, octaves=w.octaves, stream=stream)
        if ev:
            ev[3].record(stream)

    def run_steps(k, events=None):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(k):
            step(events[i] if events else None)
        if pipe is not None:
            pipe.drain()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / k

    run_steps(args.warmup)
    # timing events on every step, or (pipelined) on every tenth: an event
    # pair on each stream at every step takes away part of the overlap
    every = 10 if pipe is not None and args.steps >= 10 else 1
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] if i % every == 0 else None
              for i in range(args.steps)]
    ms_per_step = run_steps(args.steps, events)
    events = [e for e in events if e]
    serial_ms = None
    if pipe is not None:
        # the same steps one after another on one stream (not the measured value)
        workspace = pipe.workspaces[0]

        def serial_step():
            if attrs is not None:
                expr.evaluate_root_attributes(asset, w.params, exprs, maps, attrs, stream=stream)
            nx.prepare_packed(asset, w.params, workspace, stream=stream)
            nx.generate_packed_prepared(asset, w.params, out, workspace, octaves=w.octaves, stream=stream,
                                        attributes=attrs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            serial_step()
        torch.cuda.synchronize()
        serial_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    prep_ms = [e[0].elapsed_time(e[1]) for e in events]
    kern_ms = [e[2].elapsed_time(e[3]) for e in events]  # the strand kernel (expr: + guide prep and order)
    kern_ms_sorted = sorted(kern_ms)
    kernel_ms = sum(kern_ms) / len(kern_ms)
    points = N * Cv
    alg_bytes = points * 16 + N * 4 + blob.nbytes
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    traffic = valu = None
    if os.path.exists(PMC_SUMMARY) and args.workload == "plane":
        with open(PMC_SUMMARY) as f:
            pmc = json.load(f)
        traffic = pmc.get("hbm_bytes_per_launch")
        # VALU issue: wave64 VALU instructions per launch (profiled) over the kernel time
        insts = pmc.get("sq", {}).get("SQ_INSTS_VALU")
        if insts:
            valu = valu_issue(insts, kernel_ms, "instructions_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": ("strand_kernel<true, %d>" % (w.octaves.count if w.octaves else 1)
                           if attrs is None or pipe is not None
                           else "guide_prep_kernel + strand order + strand_kernel<true, 1>"),
                "kernel_ms_mean": round(kernel_ms, 4),
                "kernel_ms_median": round(kern_ms_sorted[len(kern_ms) // 2], 4),
                "kernel_ms_p90": round(kern_ms_sorted[min(len(kern_ms) - 1, int(0.9 * len(kern_ms)))], 4),
                ("prep_and_order_ms_mean" if attrs is None else
                 "attributes_prep_and_order_ms_mean" if pipe is not None else "attributes_ms_mean"):
                    round(sum(prep_ms) / len(prep_ms), 4),
                "algorithmic_bytes_per_launch": alg_bytes, "valu": valu}
    if serial_ms is not None:
        roofline["timed_launches"] = f"every {every}th step of the timed region"
        roofline["serial_ms_per_step"] = round(serial_ms, 4)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(blob, w, args.cpu_seconds, baseline_threads())
    # the metric's surrogate clause: the learned-surrogate workload beside the
    # exact runtime at the stated threshold (its own timed steps)
    sur = None
    if args.workload == "plane" and not args.no_surrogate:
        del out, ws
        sur = surrogate_record(args)
    line = {"metric": METRIC, "value": points / (ms_per_step * 1e-3), "unit": "points/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": w.name, "triangles": asset.header.triangle_count,
                       "guides": asset.header.guide_count, "strands": N, "cvs_per_strand": Cv,
                       "stencil_k": 8 if args.workload == "plane" else 4,
                       "octaves": ([[w.octaves.amplitude[i], w.octaves.frequency[i]] for i in range(w.octaves.count)]
                                   if w.octaves else [[w.params.noise_amplitude, w.params.noise_frequency]]),
                       "attribute_programs": len(w.info.get("programs", [])), "parallelism": "single-gpu",
                       "schedule": (("pipelined: each step's %s on a side stream overlaps the previous "
                                     "step's strand kernel" % ("attributes and preparation" if attrs is not None
                                                               else "preparation"))
                                    if pipe is not None else "serial")},
            "roofline": roofline, "cpu_baseline": cpu, "surrogate": sur}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
