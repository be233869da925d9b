"""Tile autotuner (project-owned; the reference has fixed 32x4 blocks,
src/include/bp.h:8-9).

For a batch it searches, per label-capacity class (16/32/64/128), the number
of slots in one work item, and the number of grid-stride blocks per compute
unit, timing whole graph replays on the GPU with events on the launch
stream.  Every candidate computes the same results (the tile only changes
how work is split), so the search is free to pick the fastest.  Results are
cached in a JSON file keyed by architecture, workload signature and label
class.
"""
from __future__ import annotations

import json
import os
import statistics

PKG = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(PKG, "tuning", "gfx950.json")

SLOT_CHOICES = (64, 128, 256, 512, 1024)
BPC_CHOICES = (2, 3, 4, 6, 8)


def _time_runs(torch, fn, reps: int) -> float:
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    return statistics.median(times)


def signature(batch) -> str:
    """Workload key: instance count, label-class histogram, total node-labels (log2 bucket)."""
    import math
    hist = {c: 0 for c in (16, 32, 64, 128)}
    tot = 0
    for d in batch.descs:
        cap = 16 if d.num_labels <= 16 else 32 if d.num_labels <= 32 else 64 if d.num_labels <= 64 else 128
        hist[cap] += 1
        tot += d.width * d.height * d.num_labels
    return f"n{batch.n}-L{hist[16]}.{hist[32]}.{hist[64]}.{hist[128]}-s{int(math.log2(max(tot, 1)))}"


def load_cache(path: str = CACHE) -> dict:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def apply_cached(batch, path: str = CACHE) -> bool:
    entry = load_cache(path).get("batch", {}).get(signature(batch))
    if not entry:
        return False
    batch.set_tiles(entry["slots_per_item"], entry["blocks_per_cu"])
    return True


def tune_batch(batch, data_ptrs, label_ptrs, reps: int = 3, path: str | None = CACHE, log=print) -> dict:
    """Coordinate search over (slots per item per class) then blocks per CU."""
    import torch
    stream = torch.cuda.current_stream().cuda_stream

    def run():
        batch.run(data_ptrs, label_ptrs, stream)

    slots = [256, 256, 256, 256]
    bpc = 4
    results = []

    def measure(sl, b):
        batch.set_tiles(sl, b)
        run()  # capture + warm-up
        torch.cuda.synchronize()
        t = _time_runs(torch, run, reps)
        results.append({"slots_per_item": list(sl), "blocks_per_cu": b, "ms": t})
        log(f"  tiles {sl} bpc {b}: {t:.3f} ms")
        return t

    best = measure(slots, bpc)
    for k in range(4):
        for c in SLOT_CHOICES:
            if c == slots[k]:
                continue
            trial = list(slots)
            trial[k] = c
            t = measure(trial, bpc)
            if t < best:
                best, slots = t, trial
    for b in BPC_CHOICES:
        if b == bpc:
            continue
        t = measure(slots, b)
        if t < best:
            best, bpc = t, b
    batch.set_tiles(slots, bpc)
    entry = {"slots_per_item": slots, "blocks_per_cu": bpc, "ms": best, "candidates": results}
    if path:
        cache = load_cache(path)
        cache.setdefault("batch", {})[signature(batch)] = {k: entry[k] for k in ("slots_per_item", "blocks_per_cu", "ms")}
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return entry
