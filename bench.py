"""Benchmark: hierarchical min-sum BP on one MI355X.

Workload (default): 1920x1080 grid, 128 labels, FP32 costs from the seeded
synthetic stereo generator (seed 2, 64 rectangles, disparities 0..127,
truncation 40), truncated-linear smoothness lambda=8 tau=4
(useSmoothnessTruncLinear(8, 32)), 5 pyramid levels x 10 red-black
half-sweeps per level, reference (mean) normalisation.  One step is one
full decodeHbp on device-resident costs: relayout + pyramid + all sweeps +
labelling.  Metric: node-label message updates per second (one update = one
node x one label with its 4 outgoing messages), counted over every
half-sweep of every level.

Prints ONE JSON line.  Only one GPU: --gpus > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E datasheet peak
BYTES_PER_UPDATE = 36     # D (4) + 4 incoming (16) + 4 outgoing (16) FP32, per node-label
                          # (34 with 16-bit data costs)


def cpu_model() -> dict:
    """Host CPU description from /proc/cpuinfo: model name, family/model/
    stepping, the logical CPUs this process may run on and their MHz."""
    info = {"model_name": platform.processor() or "unknown"}
    try:
        with open("/proc/cpuinfo") as f:
            first = {}
            for line in f:
                if not line.strip():
                    if first:
                        break
                    continue
                k, _, v = line.partition(":")
                first[k.strip()] = v.strip()
        info["model_name"] = first.get("model name", info["model_name"])
        for key in ("vendor_id", "cpu family", "model", "stepping", "cpu MHz", "cache size"):
            if key in first:
                info[key.replace(" ", "_")] = first[key]
    except OSError:
        pass
    try:
        info["usable_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["usable_cpus"] = os.cpu_count()
    info["os_cpu_count"] = os.cpu_count()
    return info


def _hbp_updates(W: int, H: int, L: int, levels: int, iters: int) -> float:
    """Node-label updates of decodeHbp: interior nodes of each half-sweep's
    colour at every level (ceil halving), times L."""
    u, w, h = 0.0, W, H
    for t in range(levels):
        if t:
            w, h = (w + 1) // 2, (h + 1) // 2
        u += _interior_updates(w, h, L, iters)
    return u


def cpu_baseline(threads_all: int, gpu_labels=None, runs_all: int = 3, log=None) -> dict:
    """The oracle (plain-C restatement of the reference CPU BP, src/bp.cc:86-180,
    built -O2 -m64 as the reference Makefile builds, Makefile:57,75-76) on the
    benchmark workload itself: 1920x1080x128, 5 levels x 10 half-sweeps, mean
    normalisation.  One run on one thread (the reference is single-threaded)
    and the median of `runs_all` runs on `threads_all` threads (OpenMP across
    the nodes of one colour, bit-identical to one thread).  The labels of every
    run are compared with the GPU's."""
    from oracle import oracle as O
    from gridinfcuda_amd import _lib as G
    log = log or (lambda m: print(m, file=sys.stderr, flush=True))
    W, H, L = 1920, 1080, 128
    data = G.gen_stereo(2, W, H, L, 64, 127, 40)["cost_f32"]
    table = O.table_trunc_linear(L, 8, 32)
    U = _hbp_updates(W, H, L, 5, 10)
    times, same = {1: [], threads_all: []}, True
    plan = [1] + [threads_all] * runs_all
    for k, th in enumerate(plan):
        log(f"cpu baseline: run {k + 1}/{len(plan)} on {th} thread(s)")
        t0 = time.perf_counter()
        lab = O.hbp(data, table, 2, 5, 10, norm=0, threads=th)
        dt = time.perf_counter() - t0
        times[th].append(dt)
        log(f"cpu baseline: {dt:.1f} s")
        if gpu_labels is not None:
            same = same and bool(np.array_equal(lab, gpu_labels))
    t1 = float(np.median(times[1]))
    ta = float(np.median(times[threads_all]))
    return {
        "value": U / t1 / 1e9, "unit": "Gnode-label msg updates/s", "cores": 1, "kind": "port",
        "sample": f"the whole workload: 1920x1080x128 HBP 5 levels x 10 half-sweeps, {U:.4g} node-label updates, "
                  f"oracle built -O2 -m64; 1 run on 1 thread, median of {runs_all} runs on {threads_all} threads",
        "seconds_1_thread": [round(t, 2) for t in times[1]],
        "value_all_cores": U / ta / 1e9, "cores_all": threads_all,
        "seconds_all_cores": [round(t, 2) for t in times[threads_all]],
        "labels_equal_gpu": same if gpu_labels is not None else None,
        "cpu": cpu_model(),
    }


def _interior_updates(W: int, H: int, L: int, iters: int) -> float:
    """Node-label updates of plain BP (one level): interior nodes of the
    half-sweep's colour, times L, summed over the half-sweeps."""
    if W < 3 or H < 3:
        return 0.0
    ys, xs = np.mgrid[1:H - 1, 1:W - 1]
    par = (xs + ys) & 1
    odd = float(par.sum())
    even = float(par.size) - odd
    # half-sweep i updates (x + y + i) odd: i even -> odd parity, i odd -> even
    return L * (odd * ((iters + 1) // 2) + even * (iters // 2))


def bench_batched(args) -> dict:
    """Batched plain BP: 64 instances (widths log-uniform 160..1280, heights
    120..720, labels from {16,32,64,128}, Potts w=15 / TL lambda=8 tau=4
    alternating), 20 half-sweeps each, one grid-stride work-list launch per half-sweep and capacity,
    the whole decode replayed as one HIP graph; tiles from the autotuner."""
    import torch
    from gridinfcuda_amd import _lib as G
    from gridinfcuda_amd import batch as B
    from gridinfcuda_amd import autotune
    insts = B.mixed_instances(64, 3000)
    datas = [torch.from_numpy(B.instance_costs(x)).cuda() for x in insts]
    labels = [torch.zeros((x["H"], x["W"]), dtype=torch.int32, device="cuda") for x in insts]
    bt = B.make_batch(insts, iters=20)
    dp = [t.data_ptr() for t in datas]
    lp = [t.data_ptr() for t in labels]
    sp = torch.cuda.current_stream().cuda_stream
    tuned = "cache"
    if args.retune or not autotune.apply_cached(bt):
        entry = autotune.tune_batch(bt, dp, lp, reps=3, path=autotune.CACHE if args.retune else None,
                                    log=lambda m: print(m, file=sys.stderr))
        tuned = "searched now"
    if args.copy_mib:  # profiler counter calibration: one copy of known size
        src = torch.ones(args.copy_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
    for _ in range(max(1, args.warmup)):
        bt.run(dp, lp, sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bt.run(dp, lp, sp)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / args.steps
    U = bt.updates_per_run()
    # HBM bytes per graph replay, measured with rocprofv3 FETCH_SIZE /
    # WRITE_SIZE (scaled by a 1 GiB copy) on this workload; see the profile summary.
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "batched_64mrf_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes_per_replay"]
    cpu = None
    if not args.no_cpu:
        # The oracle, one thread, on the batch's first instances up to ~1.5e8
        # node-label updates (a bounded sample; the mean normalisation the
        # batch runs), with their labels checked against the replay's.
        from oracle import oracle as O
        done, secs, same, names = 0.0, 0.0, True, []
        for k, x in enumerate(insts):
            u = _interior_updates(x["W"], x["H"], x["L"], 20)
            if done and done + u > 1.5e8:
                break
            d = B.instance_costs(x)
            t0 = time.perf_counter()
            ref = O.hbp(d, x["table"], x["model"], 1, 20, norm=0)
            secs += time.perf_counter() - t0
            done += u
            same = same and bool(np.array_equal(ref, labels[k].cpu().numpy()))
            names.append(f"{x['W']}x{x['H']}x{x['L']}")
        cpu = {"value": done / secs / 1e9, "unit": "Gnode-label msg updates/s", "cores": 1, "kind": "port",
               "sample": f"the oracle on the batch's first {len(names)} instances ({', '.join(names)}), "
                         f"20 half-sweeps each, {done:.4g} node-label updates, one run",
               "labels_equal_gpu": same, "seconds": round(secs, 3), "cpu_model": cpu_model()}
    value = U / (ms / 1000.0) / 1e9
    # the generator's costs are integers in [0, 40]: every instance keeps its
    # data costs in 8-bit storage (33 B per update)
    bpu = BYTES_PER_UPDATE - 3
    achieved = bpu * U / (ms / 1000.0) / 1e9
    return {
        "metric": "Gnode-label msg updates/s", "value": round(value, 3), "unit": "Gnode-label msg updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded stereo generator, seeds 3000..3063)",
        "config": {"workload": "batched plain BP, 64 mixed MRFs, 20 half-sweeps, one graph replay per step",
                   "instances": len(insts), "node_label_updates_per_step": U,
                   "label_hist": {str(c): sum(1 for x in insts if x["L"] == c) for c in B.LABEL_CLASSES},
                   "tiles": {"slots_per_item": list(bt.tiles[0]), "blocks_per_cu": bt.tiles[1], "source": tuned}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "whole graph (k_sweep_batch16 + relayouts), 33 B per node-label update "
                               "(8-bit data costs); traffic per replay"},
        "cpu_baseline": cpu,
    }


def bench_config1(args) -> dict:
    """The reference's CPU-sized case: 128x96 grid, 16 labels, FP32 integer
    costs (seed 1, 8 rectangles, disparities 0..15, truncation 20),
    V = 10*min(|a-b|, 2), decodeBp with 10 half-sweeps, min normalisation.
    GPU: device-resident plan runs, median of >= 100; CPU: the oracle on one
    thread, median of >= 100 runs (the full workload, not a sample)."""
    import torch
    from gridinfcuda_amd import _lib as G
    from oracle import oracle as O
    W, H, L, iters = 128, 96, 16, 10
    data = G.gen_stereo(1, W, H, L, 8, 15, 20)["cost_f32"]
    table = O.table_trunc_linear(L, 10, 20)
    plan = G.Plan(G.make_desc(W, H, L, G.F32, G.SM_TRUNC_LINEAR, 1, iters, G.NORM_MIN), table)
    dd = torch.from_numpy(data).cuda()
    lab = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream()
    runs = max(100, args.steps)
    for _ in range(max(1, args.warmup)):
        plan.run(dd.data_ptr(), lab.data_ptr(), s.cuda_stream)
    ts = []
    for _ in range(runs):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        plan.run(dd.data_ptr(), lab.data_ptr(), s.cuda_stream)
        b.record(s)
        b.synchronize()
        ts.append(a.elapsed_time(b))
    gpu_ms = float(np.median(ts))
    U = plan.updates_per_run()
    cts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        ref = O.hbp(data, table, 2, 1, iters, norm=1)
        cts.append((time.perf_counter() - t0) * 1000.0)
    cpu_ms = float(np.median(cts))
    same = bool(np.array_equal(ref, lab.cpu().numpy()))
    return {
        "metric": "Gnode-label msg updates/s", "value": round(U / gpu_ms / 1e6, 3), "unit": "Gnode-label msg updates/s",
        "n_gpus": 1, "steps": runs, "warmup": args.warmup, "ms_per_step": round(gpu_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded stereo generator, seed 1)",
        "config": {"workload": "plain BP 128x96, 16 labels, V=10*min(|a-b|,2), 10 half-sweeps, min normalisation",
                   "node_label_updates_per_step": U, "labels_equal_oracle": same},
        "roofline": None,
        "cpu_baseline": {"value": round(U / cpu_ms / 1e6, 5), "unit": "Gnode-label msg updates/s", "cores": 1,
                         "kind": "port", "sample": f"the full workload, median of {runs} runs",
                         "ms_per_step": round(cpu_ms, 4), "cpu_model": cpu_model()},
    }


def _stereo_table(O, L, lam, tau, dtype):
    return O.table_trunc_linear(L, lam, lam * tau, dtype)


def bench_icm(args) -> dict:
    """ICM (src/icm.cc:5-62, ginf_icm) on the 1920x1080x128 stereo costs of
    the headline workload (seed 2, TL lambda=8 tau=4, FP32), from the
    decodeMaxPrior labelling, 10 iterations, device-resident, HIP events on
    the launch stream (median of the timed runs).  Metric: node-label
    evaluations/s (interior nodes x labels x iterations / time).  The
    synchronous call's first iteration reads the costs in T and writes an
    8-bit copy of them, the later ones read the copy (1 B per node-label plus
    8 B of labels per node).  The asynchronous mode (the reference's in-place row-major
    order) runs as a wavefront over 32 x 32 tiles (one launch per tile
    anti-diagonal and pipelined iteration) and is latency-bound.  CPU:
    the oracle, one thread, on the same recipe at 960x540 (a quarter of the
    nodes), whose labels the GPU reproduces."""
    import torch
    from gridinfcuda_amd import _lib as G
    from oracle import oracle as O
    W, H, L, N = 1920, 1080, 128, 10
    s = torch.cuda.current_stream()

    def gpu_run(Wc, Hc, mode, runs):
        data = torch.from_numpy(G.gen_stereo(2, Wc, Hc, L, 64, L - 1, 40)["cost_f32"]).cuda()
        table = torch.from_numpy(_stereo_table(O, L, 8, 4, np.float32)).cuda()
        init = torch.zeros((Hc, Wc), dtype=torch.int32, device="cuda")
        out = torch.zeros_like(init)
        G.check(G.lib().ginf_max_prior(Wc, Hc, L, G.F32, ctypes.c_void_p(data.data_ptr()),
                                       ctypes.c_void_p(init.data_ptr()), ctypes.c_void_p(s.cuda_stream)))
        go = lambda: G.icm(Wc, Hc, L, G.F32, data.data_ptr(), table.data_ptr(), init.data_ptr(), N, mode,
                           out.data_ptr(), s.cuda_stream)
        for _ in range(max(1, args.warmup)):
            go()
        ts = []
        for _ in range(runs):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(s)
            go()
            b.record(s)
            b.synchronize()
            ts.append(a.elapsed_time(b))
        return float(np.median(ts)), data, init, out

    runs = max(5, args.steps)
    evals = (W - 2) * (H - 2) * L * N
    if args.copy_mib:  # profiler counter calibration: one copy of known size
        src = torch.ones(args.copy_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
    sync_ms, data, init, _ = gpu_run(W, H, G.ICM_SYNC, runs)
    async_ms, _, _, _ = gpu_run(W, H, G.ICM_ASYNC, max(3, runs // 2))
    del data, init
    # Algorithmic bytes of one synchronous call: the first iteration reads the
    # costs in T (4 B per interior node-label), writes their 8-bit copy (1 B)
    # and the labels (8 B per node); the other N - 1 read the copy (1 B per
    # node-label + 8 B per node).  The stereo costs are integers in [0, 40],
    # so the copy holds them all.
    nin = (W - 2) * (H - 2)
    algo = nin * (4 * L + L + 8) + (N - 1) * nin * (L + 8)
    # HBM bytes per synchronous call, measured with rocprofv3 FETCH_SIZE /
    # WRITE_SIZE (scaled by a 1 GiB copy) on this workload; see the profile summary.
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "icm_sync_1920x1080x128_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes_per_call"]
    # CPU sample and its parity check
    Ws, Hs = 960, 540
    _, sd, sinit, sout = gpu_run(Ws, Hs, G.ICM_SYNC, 1)
    cdata = G.gen_stereo(2, Ws, Hs, L, 64, L - 1, 40)["cost_f32"]
    ctab = _stereo_table(O, L, 8, 4, np.float32)
    cinit = O.max_prior(cdata)
    t0 = time.perf_counter()
    ref = O.icm(cdata, ctab, cinit, N, sync=True)
    cpu_s = time.perf_counter() - t0
    same = bool(np.array_equal(ref, sout.cpu().numpy())) and bool(np.array_equal(cinit, sinit.cpu().numpy()))
    cpu_evals = (Ws - 2) * (Hs - 2) * L * N
    return {
        "metric": "Gnode-label ICM evaluations/s", "value": round(evals / sync_ms / 1e6, 3),
        "unit": "Gnode-label evaluations/s", "n_gpus": 1, "steps": runs, "warmup": args.warmup,
        "ms_per_step": round(sync_ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded stereo generator, seed 2)",
        "config": {"workload": "ICM 1920x1080x128 TL lambda=8 tau=4, from decodeMaxPrior, 10 iterations",
                   "sync_ms": round(sync_ms, 4), "async_ms": round(async_ms, 4),
                   # asynchronous ICM with the table in LDS (L <= 128): one launch per 32 x 32 tile
                   # anti-diagonal and pipelined iteration
                   "async_launches": (-(-(W - 2) // 32)) + (-(-(H - 2) // 32)) - 1 + 2 * (N - 1),
                   "labels_equal_oracle_at_960x540": same},
        "roofline": {"bound": "hbm", "achieved": round(algo / sync_ms / 1e6, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(algo / sync_ms / 1e6 / 8000.0, 4), "traffic": traffic,
                     "kernel": "synchronous ICM call: the first of 10 launches reads the costs in T and writes "
                               "their 8-bit copy, the others read the copy (1 B per node-label + 8 B per node); "
                               "LDS- and VALU-limited since the copy; traffic per call"},
        "cpu_baseline": {"value": round(cpu_evals / cpu_s / 1e9, 5), "unit": "Gnode-label evaluations/s", "cores": 1,
                         "kind": "port", "sample": "the same recipe at 960x540, 10 synchronous iterations, one run",
                         "seconds": round(cpu_s, 3), "cpu_model": cpu_model()},
    }


def bench_alpha_beta(args) -> dict:
    """Alpha-beta swap (src/graphcut.cc:194-279, ginf_alpha_beta) on int
    stereo costs at the Tsukuba size: 384x288, 16 labels (seed 1, 8
    rectangles, disparities 0..15, truncation 20), V = 10*min(|a-b|, 2), from
    the decodeMaxPrior labelling, one cycle over all 120 label pairs;
    device-resident, median of the timed runs.  Each pair is a maximum
    preflow by parallel push-relabel (atomics, global relabels): latency- and
    atomic-bound, so no roofline is claimed.  CPU: the oracle (each pair
    solved by BFS augmenting paths), one thread, the full workload once; both
    give the same labels (the sink-reachable set of a maximum flow is
    canonical)."""
    import torch
    from gridinfcuda_amd import _lib as G
    from oracle import oracle as O
    W, H, L = 384, 288, 16
    cost = G.gen_stereo(1, W, H, L, 8, 15, 20, want=("cost_i32",))["cost_i32"]
    table = _stereo_table(O, L, 10, 2, np.int32)
    s = torch.cuda.current_stream()
    dd = torch.from_numpy(cost).cuda()
    td = torch.from_numpy(table).cuda()
    init = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    out = torch.zeros_like(init)
    G.check(G.lib().ginf_max_prior(W, H, L, G.I32, ctypes.c_void_p(dd.data_ptr()), ctypes.c_void_p(init.data_ptr()),
                                   ctypes.c_void_p(s.cuda_stream)))
    go = lambda: G.alpha_beta(W, H, L, dd.data_ptr(), td.data_ptr(), init.data_ptr(), 1, out.data_ptr(), s.cuda_stream)
    for _ in range(max(1, args.warmup)):
        go()
    runs = max(5, args.steps)
    ts, moves = [], 0
    for _ in range(runs):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        moves = go()
        b.record(s)
        b.synchronize()
        ts.append(a.elapsed_time(b))
    gpu_ms = float(np.median(ts))
    t0 = time.perf_counter()
    ref, ref_moves = O.alpha_beta(cost, table, O.max_prior(cost), 1)
    cpu_s = time.perf_counter() - t0
    same = bool(np.array_equal(ref, out.cpu().numpy())) and ref_moves == moves
    return {
        "metric": "ms per alpha-beta cycle", "value": round(gpu_ms, 3), "unit": "ms", "n_gpus": 1, "steps": runs,
        "warmup": args.warmup, "ms_per_step": round(gpu_ms, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded stereo generator, seed 1)",
        "config": {"workload": "alpha-beta swap 384x288x16, V=10*min(|a-b|,2), one cycle over 120 label pairs",
                   "moves_kept": moves, "labels_and_moves_equal_oracle": same},
        "roofline": None,
        "cpu_baseline": {"value": round(cpu_s * 1000.0, 1), "unit": "ms", "cores": 1, "kind": "port",
                         "sample": "the full workload, one run", "cpu_model": cpu_model()},
    }


def bench_stereo(args) -> dict:
    """The reference's stereo example (examples/stereo/stereo.cc:25-51) on the
    Tsukuba pair it ships: 384x288, 16 labels, Grid<int> costs
    min(|L - R(x-k)|, 20), TL (10, 20), decodeBp with 30 half-sweeps
    (reference mean normalisation, integer arithmetic).  GPU: device-resident
    plan runs (median of >= 100) and the synchronous host entry the C++
    example's decodeBp goes through (median of 20, copies included); CPU: the
    oracle on one thread, median of 5 runs of the full workload."""
    import torch
    from gridinfcuda_amd import _lib as G
    from gridinfcuda_amd import stereo as S
    from oracle import oracle as O
    gold = os.path.join(ROOT, "tests", "golden")
    left, right = S.load_gray(os.path.join(gold, "tsukuba_left.png")), S.load_gray(os.path.join(gold, "tsukuba_right.png"))
    costs = S.example_costs(left, right)
    table = S.smoothness_table()
    L, H, W = costs.shape
    iters = S.ITERS
    plan = G.Plan(G.make_desc(W, H, L, G.I32, G.SM_TRUNC_LINEAR, 1, iters, G.NORM_MEAN), table)
    dd = torch.from_numpy(np.ascontiguousarray(costs)).cuda()
    lab = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream()
    for _ in range(max(1, args.warmup)):
        plan.run(dd.data_ptr(), lab.data_ptr(), s.cuda_stream)
    runs = max(100, args.steps)
    ts = []
    for _ in range(runs):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        plan.run(dd.data_ptr(), lab.data_ptr(), s.cuda_stream)
        b.record(s)
        b.synchronize()
        ts.append(a.elapsed_time(b))
    gpu_ms = float(np.median(ts))
    host = []
    for _ in range(20):
        t0 = time.perf_counter()
        hl = G.hbp_host(costs, table, G.SM_TRUNC_LINEAR, 1, iters)
        host.append((time.perf_counter() - t0) * 1000.0)
    cts = []
    for _ in range(5):
        t0 = time.perf_counter()
        ref = O.hbp(costs, table, G.SM_TRUNC_LINEAR, 1, iters)
        cts.append((time.perf_counter() - t0) * 1000.0)
    cpu_ms = float(np.median(cts))
    U = plan.updates_per_run()
    same = bool(np.array_equal(ref, lab.cpu().numpy())) and bool(np.array_equal(ref, hl))
    return {
        "metric": "Gnode-label msg updates/s", "value": round(U / gpu_ms / 1e6, 3), "unit": "Gnode-label msg updates/s",
        "n_gpus": 1, "steps": runs, "warmup": args.warmup, "ms_per_step": round(gpu_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "Tsukuba stereo pair shipped with the reference example",
        "config": {"workload": "stereo example 384x288x16, Grid<int>, TL (10, 20), decodeBp 30 half-sweeps",
                   "node_label_updates_per_step": U, "host_entry_ms": round(float(np.median(host)), 4),
                   "labels_equal_oracle": same},
        "roofline": None,
        "cpu_baseline": {"value": round(U / cpu_ms / 1e6, 5), "unit": "Gnode-label msg updates/s", "cores": 1,
                         "kind": "port", "sample": "the full workload, median of 5 runs", "ms_per_step": round(cpu_ms, 3),
                         "cpu_model": cpu_model()},
    }


def hbm_probe(torch, G, mib_per_stream: int = 512) -> dict:
    """Best bandwidth of the library's streaming probe (ginf_stream_probe):
    5 read + 4 written float4 streams (the half-sweep's stream shape) and a
    1 + 1 copy, each over a few grid sizes; HIP events on the launch stream."""
    n = (mib_per_stream << 20) // 16
    inp = torch.ones(5 * n * 4, device="cuda")
    out = torch.empty(4 * n * 4, device="cuda")
    s = torch.cuda.current_stream()
    res = {}
    for (r, w) in ((5, 4), (1, 1)):
        best = 0.0
        for blocks in (16384, 65536, 262144):
            def go():
                G.check(G.lib().ginf_stream_probe(r, w, ctypes.c_void_p(inp.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                                  n, blocks, ctypes.c_void_p(s.cuda_stream)))
            go()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(s)
            for _ in range(10):
                go()
            b.record(s)
            b.synchronize()
            ms = a.elapsed_time(b) / 10
            best = max(best, (r + w) * n * 16 / ms / 1e6)
        res[f"{r}r{w}w"] = round(best, 1)
    del inp, out
    return res


def bench_warmstart(args) -> dict:
    """Learned warm start, 640x480, 64 labels, TL lambda=10 tau=3: per seeded
    evaluation instance (seeds 2000000...), plain HBP 5x10 (the energy reference) against
    each seeded schedule, with the committed network and with zero seeds.
    value = time of the cheapest schedule that reaches plain HBP's energy
    within 0.1 % on every instance, divided by plain HBP's time."""
    import torch
    from gridinfcuda_amd import warmstart as WS
    wl = WS.Workload()
    net = WS.ConvNet.load(WS.default_model_path(wl))
    seeds = list(range(2_000_000, 2_000_000 + args.eval_instances))
    log = lambda m: print(m, file=sys.stderr)
    rec_net = WS.evaluate(wl, net, seeds, reps=args.steps, log=log)
    rec_zero = WS.evaluate(wl, None, seeds, reps=args.steps, log=log)
    rows = {"net": WS.summarize(rec_net), "zero_seed": WS.summarize(rec_zero)}
    plain_ms = float(np.median([r["plain_ms"] for r in rec_net + rec_zero]))

    def best(rs):
        ok = [r for r in rs if r["within_0.1pct"] >= 1.0]
        return min(ok, key=lambda r: r["time_ratio"]) if ok else None
    b_net, b_zero = best(rows["net"]), best(rows["zero_seed"])
    return {
        "metric": "ms to plain-BP energy, warm-start vs plain", "value": b_net["time_ratio"] if b_net else None,
        "unit": "x plain HBP time", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": None,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32 BP, bf16 network",
        "data": f"synthetic (seeded stereo generator, seeds {seeds[0]}..{seeds[-1]})",
        "config": {"workload": "learned warm start 640x480x64 TL lambda=10 tau=3 vs plain HBP 5 levels x 10",
                   "instances": len(seeds), "target": "energy within 0.1% of plain HBP on every instance"},
        "best_with_net": b_net, "best_zero_seed": b_zero, "plain_ms": round(plain_ms, 4), "schedules": rows,
        "roofline": None, "cpu_baseline": None,
    }


def bench_warmstart_batch(args) -> dict:
    """Batched learned warm start: 256 instances (640x480x64 / 1280x720x96,
    TL and Potts), batched network inference, seeded batched finest-level BP
    of K half-sweeps, fallback to plain HBP above 0.1 % energy excess."""
    from gridinfcuda_amd import warmstart_batch as WB
    ks = [int(v) for v in args.batch_iters.split(",")]
    res = WB.evaluate_batched(args.eval_instances, iters=ks, reps=max(1, args.steps),
                              log=lambda m: print(m, file=sys.stderr))
    s = min(res["summary"], key=lambda r: r["time_ratio"])
    return {
        "metric": "ms to plain-BP energy, warm-start vs plain", "value": s["time_ratio"], "unit": "x plain HBP time",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": s["warm_total_ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32 BP, bf16 network",
        "data": f"synthetic (seeded stereo generator, seeds 5000..{5000 + args.eval_instances - 1})",
        "config": {"workload": "batched learned warm start, 640x480x64 and 1280x720x96, TL and Potts, "
                               "fallback to plain HBP 5x10 above 0.1% energy excess",
                   "instances": args.eval_instances, "chunk": res["chunk"], "tiles": res["tiles"]},
        "fallback_fraction": s["fallback_fraction"], "summary": res["summary"],
        "roofline": None, "cpu_baseline": None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--labels", type=int, default=128)
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per node (0 = tuned default)")
    ap.add_argument("--norm", type=int, default=0, help="0 mean (reference), 1 min")
    ap.add_argument("--kernel", type=int, default=0, help="0 auto, 1 chunked, 2 quad (per-direction lanes)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-api", action="store_true", help="skip the host-entry (PCIe-inclusive) timing")
    ap.add_argument("--no-probe", action="store_true", help="skip the HBM streaming probe")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload: skip the warm-start and batched figures added to the line")
    ap.add_argument("--pairs", type=int, default=None, choices=[0, 1],
                    help="fused pairs of half-sweeps: 1 on, 0 off (default: the plan's choice, on up to 32 labels)")
    ap.add_argument("--data32", action="store_true",
                    help="keep 32-bit data costs (no 16- or 8-bit storage even when the costs fit)")
    ap.add_argument("--data16", action="store_true",
                    help="at most 16-bit data costs (no 8-bit storage of levels 0 and 1)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads for the all-cores CPU figure")
    ap.add_argument("--workload", default="hbp", choices=["hbp", "batched", "warmstart", "warmstart-batch", "config1",
                                                          "icm", "alpha-beta", "stereo"],
                    help="hbp: 1920x1080x128 hierarchical BP (default); batched: 64 mixed MRFs; "
                         "warmstart: 640x480x64 learned warm start vs plain HBP; "
                         "warmstart-batch: 256-instance batched warm start with fallback; "
                         "config1: 128x96x16 plain BP, GPU and one CPU core; "
                         "icm: ICM on the default workload's costs; alpha-beta: 384x288x16 alpha-beta swap; "
                         "stereo: the reference example on the Tsukuba pair")
    ap.add_argument("--eval-instances", type=int, default=None,
                    help="warm-start workloads: instances evaluated (default 8 / 256)")
    ap.add_argument("--batch-iters", default="16", help="warmstart-batch: finest-level half-sweeps K (list)")
    ap.add_argument("--retune", action="store_true", help="batched: rerun the tile search and rewrite the cache")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (profiler counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is designed, not built)"}))
        return 2

    import torch
    from gridinfcuda_amd import _lib as G

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        return 1
    if args.workload == "batched":
        print(json.dumps(bench_batched(args)))
        return 0
    if args.workload == "config1":
        print(json.dumps(bench_config1(args)))
        return 0
    if args.workload == "warmstart":
        args.eval_instances = args.eval_instances or 8
        print(json.dumps(bench_warmstart(args)))
        return 0
    if args.workload == "warmstart-batch":
        args.eval_instances = args.eval_instances or 256
        print(json.dumps(bench_warmstart_batch(args)))
        return 0
    if args.workload == "icm":
        print(json.dumps(bench_icm(args)))
        return 0
    if args.workload == "alpha-beta":
        print(json.dumps(bench_alpha_beta(args)))
        return 0
    if args.workload == "stereo":
        print(json.dumps(bench_stereo(args)))
        return 0
    W, H, L = args.width, args.height, args.labels
    gen = G.gen_stereo(2, W, H, L, 64, 127, 40)["cost_f32"]
    table = np.minimum(8 * np.abs(np.arange(L)[:, None] - np.arange(L)[None, :]), 32).astype(np.float32)
    desc = G.make_desc(W, H, L, G.F32, G.SM_TRUNC_LINEAR, args.levels, args.iters, args.norm, args.lanes,
                       args.kernel)
    plan = G.Plan(desc, table)
    if args.data32:
        plan.set_compact_data(False)
    elif args.data16:
        plan.set_compact_data(2)
    if args.pairs is not None:
        plan.set_pairs(bool(args.pairs))
    data = torch.from_numpy(gen).cuda()
    del gen
    labels = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    if args.copy_mib:
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        plan.run(data.data_ptr(), labels.data_ptr(), sp)
    torch.cuda.synchronize()
    plan.sweep_stats(reset=True)

    # The per-launch HIP events of the roofline figure (two event records per
    # half-sweep, ~0.3 ms per decode) are taken during the last timed step
    # only, so `value` carries 1/steps of that overhead.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        if i == args.steps - 1:
            plan.set_sweep_timing(True)
        plan.run(data.data_ptr(), labels.data_ptr(), sp)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    level_updates = plan.sweep_level_updates()
    sweep_ms, launches, sweep_updates = plan.sweep_stats(reset=True)
    plan.set_sweep_timing(False)

    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    U = plan.updates_per_run()
    value = U / (ms_per_step / 1000.0) / 1e9

    # With 16-bit data costs (integer costs in [0, 65535], the plan's default for
    # the quad kernel) an update reads 2 B of data cost instead of 4, and with
    # 8-bit costs (levels 0 and 1, integers in [0, 255]) 1 B.
    form = plan.data_form()
    compact = form != 0
    data_bytes = [4 if form == 0 else 1 if (form == 2 and t <= 1) else 2 for t in range(args.levels)]
    alg_bytes = sum((BYTES_PER_UPDATE - 4 + db) * u for db, u in zip(data_bytes, level_updates))
    bytes_per_update = alg_bytes / max(1.0, sweep_updates)
    avg_launch_ms = sweep_ms / max(1, launches)
    bytes_per_launch = alg_bytes / max(1, launches)
    achieved = bytes_per_launch / (avg_launch_ms / 1000.0) / 1e9
    quad = args.kernel != 1
    pairs = plan.pairs()
    kname = ("k_sweep_pair (two half-sweeps per launch)" if pairs else
             ("k_sweep_quad16" if compact else "k_sweep_quad") if quad else "k_sweep_linear")

    # HBM bytes per steady half-sweep launch, measured with rocprofv3 FETCH_SIZE /
    # WRITE_SIZE (scaled by a 1 GiB copy) on this workload; see the profile summary.
    traffic = None
    tname = ["sweep_quad", "sweep_quad16", "sweep_quad16_u8"][form]
    tpath = os.path.join(ROOT, "profiles", f"{tname}_1920x1080x128_traffic.json")
    default_shape = (W, H, L, args.levels, args.iters, args.norm, args.kernel, args.lanes) == (1920, 1080, 128, 5, 10, 0, 0, 0) \
        and not pairs
    if default_shape and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes_per_launch"]

    # FP64 energy of the result (outside the timed region).
    tdev = torch.from_numpy(table).cuda()
    energy = G.energy_f64(W, H, L, G.F32, data.data_ptr(), tdev.data_ptr(), labels.data_ptr(), sp)

    shape = "kernel=%s lanes=%d" % (["auto", "chunked", "quad"][args.kernel], args.lanes)
    out = {
        "metric": "Gnode-label msg updates/s",
        "value": round(value, 3),
        "unit": "Gnode-label msg updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded stereo generator)",
        "config": {
            "workload": f"hierarchical min-sum BP {W}x{H}, {L} labels, TL lambda=8 tau=4, "
                        f"{args.levels} levels x {args.iters} half-sweeps, "
                        f"{'mean' if args.norm == 0 else 'min'} normalisation",
            "width": W, "height": H, "labels": L, "levels": args.levels, "iters_per_level": args.iters,
            "node_label_updates_per_step": U, "tile": shape, "fused_pairs": pairs,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": f"{kname} steady-state instance (half-sweep)",
            "bytes_per_update": round(bytes_per_update, 4),
            "data_costs": ["32-bit", "16-bit", "8-bit at levels 0-1, 16-bit above"][form],
            "avg_launch_ms": round(avg_launch_ms, 5),
            "launches_timed": launches,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "energy_f64": energy,
    }
    if not args.no_probe:
        # Practical ceiling for the sweep's stream shape, measured now on this GPU.
        probe = hbm_probe(torch, G)
        out["roofline"]["probe_gbs"] = probe
        out["roofline"]["frac_of_probe"] = round(achieved / probe["5r4w"], 4)
    if not args.no_api:
        # End to end through the synchronous host entry (ginf_hbp_host, what the
        # C++ templates call): pageable host costs in, labels out, PCIe both
        # ways.  First call on this problem and a repeat call; never `value`.
        host = G.gen_stereo(2, W, H, L, 64, 127, 40)["cost_f32"]
        api_ms = []
        for _ in range(2):
            t0 = time.perf_counter()
            hl = G.hbp_host(host, table, G.SM_TRUNC_LINEAR, args.levels, args.iters, args.norm,
                            lanes_per_node=args.lanes)
            api_ms.append((time.perf_counter() - t0) * 1000.0)
        out["api_end_to_end_ms"] = {"first_call": round(api_ms[0], 2), "repeat_call": round(api_ms[1], 2),
                                    "h2d_bytes": int(host.nbytes),
                                    "labels_equal_device_run": bool(np.array_equal(hl, labels.cpu().numpy()))}
        G.lib().ginf_hbp_host_release()
        del host
    gpu_labels = labels.cpu().numpy()
    del plan, data, labels, tdev
    torch.cuda.empty_cache()
    if not args.no_extras and default_shape:
        # The benchmark's other two figures on the same line: the learned warm
        # start (ms to plain-BP energy, warm start vs plain, 640x480x64 TL) and
        # the batched workload (64 mixed MRFs).  Each is the whole record its
        # own --workload run prints.
        log = lambda m: print(m, file=sys.stderr, flush=True)
        log("warm start: 4 instances of 640x480x64")
        ws = bench_warmstart(argparse.Namespace(**{**vars(args), "eval_instances": 4, "steps": 5}))
        out["warm_start"] = {
            "metric": ws["metric"], "unit": ws["unit"], "instances": ws["config"]["instances"],
            "best_with_net": ws["best_with_net"], "best_zero_seed": ws["best_zero_seed"],
            "time_ratio_with_net": ws["best_with_net"]["time_ratio"] if ws["best_with_net"] else None,
            "time_ratio_zero_seed": ws["best_zero_seed"]["time_ratio"] if ws["best_zero_seed"] else None,
            "plain_ms": ws["plain_ms"], "schedules": ws["schedules"]}
        torch.cuda.empty_cache()
        log("batched: 64 mixed MRFs")
        bb = bench_batched(argparse.Namespace(**{**vars(args), "steps": 5, "warmup": 2}))
        out["batched"] = {k: bb[k] for k in ("metric", "value", "unit", "ms_per_step", "config", "roofline",
                                             "cpu_baseline")}
        torch.cuda.empty_cache()
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(threads, gpu_labels if default_shape else None)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
