"""Batched learned warm start (project-owned; no reference counterpart).

Workload: 256 seeded stereo instances, alternating 640x480 with 64 labels and
1280x720 with 96 labels, truncated linear and Potts smoothness in turn
(index i: size class i % 2, smoothness (i // 2) % 2).  Each of the four
(size, smoothness) classes has its own network.

One pass over a chunk of instances:
  1. device costs from the image pairs (ginf_stereo_costs);
  2. per class, the cost volumes min-pooled to 1/4 resolution and one
     batched network forward on MFMA (ginf_conv3x3_fwd);
  3-4. one batched finest-level BP of K half-sweeps over all instances
     (ginf_batch_run: grid-stride work-list launches, HIP graph replay) whose
     first half-sweep reads each instance's prediction as its incoming
     messages (ginf_batch_set_predictions; 4x outbox upsample, the same
     results as seeding the outboxes with ginf_batch_seed);
  5. FP64 energies of the labellings (ginf_energy_f64).
Plain hierarchical BP (5 levels x 10 half-sweeps, one plan per class) is the
energy reference: an instance whose warm-started energy exceeds the plain energy by more
than 0.1 % falls back to the plain result, and its plain time is added to the
warm-start time.  The same batch seeded with zeros (no network) is the
ablation.
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import numpy as np

from . import _lib as G
from . import autotune
from . import batch as B
from .warmstart import (IN_SCALE, POOL, ConvNet, Workload, _ck, _timed, _vp, default_model_path, gen_images,
                        train)

SMALL = dict(W=640, H=480, L=64, rects=24)
LARGE = dict(W=1280, H=720, L=96, rects=48)


def class_workloads():
    """The four (size, smoothness) classes: TL lambda=10 tau=3 at 64 labels,
    TL lambda=8 tau=4 at 96 labels, Potts w=15 at both."""
    return [
        Workload(model=G.SM_TRUNC_LINEAR, lam=10, tau=3, **SMALL),
        Workload(model=G.SM_TRUNC_LINEAR, lam=8, tau=4, **LARGE),
        Workload(model=G.SM_POTTS, lam=15, tau=1, **SMALL),
        Workload(model=G.SM_POTTS, lam=15, tau=1, **LARGE),
    ]


def instance_class(i: int) -> int:
    return (i % 2) + 2 * ((i // 2) % 2)


FALLBACK_TOL = 1e-3


def fallback_flags(e_warm, e_plain, tol: float = FALLBACK_TOL):
    """An instance falls back when its warm-started energy exceeds the plain
    HBP energy by more than tol (relative)."""
    return [bool(w > p * (1.0 + tol)) for w, p in zip(e_warm, e_plain)]


class BatchedWarmStart:
    """Device state for one chunk layout: per-instance costs and labels, per
    class network inputs / predictions, the plain plans (energy reference) and a
    seeded batch of finest-level plans with K half-sweeps."""

    def __init__(self, classes, nets, chunk_classes, iters: int):
        import torch
        self.torch = torch
        self.classes = classes
        self.nets = nets
        self.cc = list(chunk_classes)
        self.n = len(self.cc)
        self.iters = iters
        self.costs, self.labels, self.left, self.right = [], [], [], []
        for c in self.cc:
            wl = classes[c]
            self.costs.append(torch.empty((wl.L, wl.H, wl.W), dtype=torch.float32, device="cuda"))
            self.labels.append(torch.zeros((wl.H, wl.W), dtype=torch.int32, device="cuda"))
            self.left.append(torch.empty((wl.H, wl.W), dtype=torch.uint8, device="cuda"))
            self.right.append(torch.empty((wl.H, wl.W), dtype=torch.uint8, device="cuda"))
        self.members = {c: [i for i, x in enumerate(self.cc) if x == c] for c in set(self.cc)}
        self.xin, self.pred = {}, {}
        for c, idx in self.members.items():
            wl = classes[c]
            self.xin[c] = torch.empty((len(idx), wl.h4, wl.w4, wl.L), dtype=torch.int16, device="cuda")
            self.pred[c] = torch.zeros((len(idx), wl.h4, wl.w4, 4 * wl.L), dtype=torch.float32, device="cuda")
        self.tables = {c: torch.from_numpy(classes[c].table).cuda() for c in set(self.cc)}
        self.plain_plans = {c: G.Plan(G.make_desc(classes[c].W, classes[c].H, classes[c].L, G.F32, classes[c].model,
                                                  classes[c].levels, classes[c].iters, G.NORM_MEAN), classes[c].table)
                            for c in set(self.cc)}
        self.plain_labels = [torch.zeros_like(t) for t in self.labels]
        descs = [G.make_desc(classes[c].W, classes[c].H, classes[c].L, G.F32, classes[c].model, 1, iters, G.NORM_MEAN)
                 for c in self.cc]
        self.batch = B.Batch(descs, [classes[c].table for c in self.cc])
        # the batch's first half-sweep reads each instance's prediction
        # (ginf_batch_set_predictions: no seeding pass); zero predictions are
        # the ablation
        preds = []
        for k, c in enumerate(self.cc):
            wl = classes[c]
            j = self.members[c].index(k)
            preds.append((self.pred[c][j].data_ptr(), wl.w4, wl.h4, wl.msg_scale))
        self.batch.set_predictions(preds, shift=2)
        self.tiles_cached = autotune.apply_cached(self.batch)

    def stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def load(self, seeds):
        """Images of the chunk's instances (host generator), costs on the device."""
        by_class = {}
        for k, (s, c) in enumerate(zip(seeds, self.cc)):
            by_class.setdefault(c, []).append((k, s))
        for c, ks in by_class.items():
            imgs = gen_images([s for _, s in ks], self.classes[c])
            for (k, _), (lft, rgt) in zip(ks, imgs):
                wl = self.classes[c]
                self.left[k].copy_(self.torch.from_numpy(lft))
                self.right[k].copy_(self.torch.from_numpy(rgt))
                _ck(G.lib().ginf_stereo_costs(_vp(self.left[k]), _vp(self.right[k]), wl.W, wl.H, wl.L, wl.trunc,
                                             _vp(self.costs[k]), ctypes.c_void_p(self.stream())))

    def warm(self, use_net: bool = True):
        """Steps 2-4: network inference per class into the predictions the
        batch reads, then one batched BP (its first half-sweep reads them)."""
        st = self.stream()
        for c, idx in self.members.items():
            wl = self.classes[c]
            if use_net:
                for j, k in enumerate(idx):
                    _ck(G.lib().ginf_minpool_costs(_vp(self.costs[k]), G.F32, wl.W, wl.H, wl.L, POOL, IN_SCALE,
                                                  _vp(self.xin[c][j]), ctypes.c_void_p(st)))
                self.nets[c].infer(self.xin[c], len(idx), wl.h4, wl.w4, self.pred[c])
            else:
                self.pred[c].zero_()
        self.batch.run([t.data_ptr() for t in self.costs], [t.data_ptr() for t in self.labels], st)

    def plain(self):
        st = self.stream()
        for k, c in enumerate(self.cc):
            self.plain_plans[c].run(self.costs[k].data_ptr(), self.plain_labels[k].data_ptr(), st)

    def plain_one(self, k):
        self.plain_plans[self.cc[k]].run(self.costs[k].data_ptr(), self.plain_labels[k].data_ptr(), self.stream())

    def finalize(self, e_warm, e_plain, tol: float = FALLBACK_TOL):
        """Fallback: instances whose warm-started energy is not within tol of
        plain HBP's take the plain labels.  Returns the fallback flags."""
        flags = fallback_flags(e_warm, e_plain, tol)
        for k, f in enumerate(flags):
            if f:
                self.labels[k].copy_(self.plain_labels[k])
        return flags

    def energies(self, labels):
        out = []
        for k, c in enumerate(self.cc):
            wl = self.classes[c]
            out.append(G.energy_f64(wl.W, wl.H, wl.L, G.F32, self.costs[k].data_ptr(), self.tables[c].data_ptr(),
                                    labels[k].data_ptr(), self.stream()))
        return out

    def close(self):
        self.batch.close()


def load_nets(classes, paths=None):
    nets = []
    for k, wl in enumerate(classes):
        p = paths[k] if paths else default_model_path(wl)
        nets.append(ConvNet.load(p))
    return nets


def evaluate_batched(n=256, base_seed=5000, chunk=64, iters=(16,), reps=3, nets=None, tune=True, log=print):
    """Run the batched warm start over n instances in chunks and compare every
    instance with plain HBP.  With tune, a chunk layout that has no cached
    tiles is tuned once (and cached) before it is timed.  Returns per-K
    summaries and per-instance records."""
    import torch
    classes = class_workloads()
    nets = nets if nets is not None else load_nets(classes)
    seeds = [base_seed + i for i in range(n)]
    per_k = {K: {"t_warm": 0.0, "t_zero": 0.0, "t_fallback": 0.0, "t_fallback_zero": 0.0, "records": []} for K in iters}
    t_plain_total = 0.0
    t_wall = time.time()
    for c0 in range(0, n, chunk):
        ids = list(range(c0, min(n, c0 + chunk)))
        cc = [instance_class(i) for i in ids]
        plain_e, plain_ms = None, None
        for K in iters:
            bws = BatchedWarmStart(classes, nets, cc, K)
            bws.load([seeds[i] for i in ids])
            if plain_e is None:
                bws.plain()
                torch.cuda.synchronize()
                plain_e = bws.energies(bws.plain_labels)
                plain_ms = [_timed(torch, lambda k=k: bws.plain_one(k), reps) for k in range(len(ids))]
                t_plain_total += _timed(torch, bws.plain, reps)
            bws.warm(True)  # graph capture + warm-up
            torch.cuda.synchronize()
            if tune and not bws.tiles_cached:
                autotune.tune_batch(bws.batch, [t.data_ptr() for t in bws.costs], [t.data_ptr() for t in bws.labels],
                                    reps=3, log=log)
                bws.warm(True)
                torch.cuda.synchronize()
            tiles = {"slots_per_item": list(bws.batch.tiles[0]), "blocks_per_cu": bws.batch.tiles[1]}
            t_w = _timed(torch, lambda: bws.warm(True), reps)
            e_w = bws.energies(bws.labels)
            t_z = _timed(torch, lambda: bws.warm(False), reps)
            e_z = bws.energies(bws.labels)
            rec = per_k[K]
            rec["t_warm"] += t_w
            rec["t_zero"] += t_z
            flags_w, flags_z = fallback_flags(e_w, plain_e), fallback_flags(e_z, plain_e)
            for j, i in enumerate(ids):
                fb, fz = flags_w[j], flags_z[j]
                rec["t_fallback"] += plain_ms[j] if fb else 0.0
                rec["t_fallback_zero"] += plain_ms[j] if fz else 0.0
                rec["records"].append({"index": i, "seed": seeds[i], "class": classes[cc[j]].key(),
                                       "plain_energy": plain_e[j], "plain_ms": plain_ms[j],
                                       "warm_energy": e_w[j], "zero_energy": e_z[j],
                                       "fallback": bool(fb), "fallback_zero": bool(fz)})
            log(f"  chunk {c0 // chunk}: K={K} warm {t_w:.2f} ms, zero-seed {t_z:.2f} ms, "
                f"fallbacks {sum(r['fallback'] for r in rec['records'][-len(ids):])}/{len(ids)} "
                f"(zero seed {sum(r['fallback_zero'] for r in rec['records'][-len(ids):])})")
            bws.close()
            del bws
            torch.cuda.empty_cache()
    summary = []
    for K in iters:
        rec = per_k[K]
        rs = rec["records"]
        ex_w = [r["warm_energy"] / r["plain_energy"] - 1 for r in rs]
        ex_z = [r["zero_energy"] / r["plain_energy"] - 1 for r in rs]
        row = {"K": K, "instances": len(rs),
               "fallback_fraction": float(np.mean([r["fallback"] for r in rs])),
               "fallback_fraction_zero_seed": float(np.mean([r["fallback_zero"] for r in rs])),
               "mean_rel_excess": float(np.mean(ex_w)), "mean_rel_excess_zero_seed": float(np.mean(ex_z)),
               "plain_ms": t_plain_total, "warm_batch_ms": rec["t_warm"], "zero_batch_ms": rec["t_zero"],
               "warm_total_ms": rec["t_warm"] + rec["t_fallback"],
               "zero_total_ms": rec["t_zero"] + rec["t_fallback_zero"]}
        row["time_ratio"] = row["warm_total_ms"] / t_plain_total
        row["time_ratio_zero_seed"] = row["zero_total_ms"] / t_plain_total
        per_class = {}
        for r in rs:
            d = per_class.setdefault(r["class"], {"n": 0, "fallback": 0, "fallback_zero": 0})
            d["n"] += 1
            d["fallback"] += int(r["fallback"])
            d["fallback_zero"] += int(r["fallback_zero"])
        row["per_class"] = per_class
        summary.append(row)
        log(f"K={K}: fallback {row['fallback_fraction']:.3f} (zero seed {row['fallback_fraction_zero_seed']:.3f}); "
            f"time with fallbacks {row['time_ratio']:.2f}x plain (zero seed {row['time_ratio_zero_seed']:.2f}x); "
            f"batch alone {row['warm_batch_ms'] / t_plain_total:.2f}x")
    return {"instances": n, "base_seed": base_seed, "chunk": chunk, "tiles": tiles, "wall_s": time.time() - t_wall,
            "summary": summary, "records": {K: per_k[K]["records"] for K in iters}}


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="batched learned warm start: train the class networks / evaluate")
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--cls", type=int, default=0, help="train: class index 0..3")
    ap.add_argument("--instances", type=int, default=6000)
    ap.add_argument("--budget", type=float, default=None)
    ap.add_argument("--seed-base", type=int, default=1_000_000)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--iters", default="16")
    ap.add_argument("--model-dir", default=None)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    classes = class_workloads()
    if a.cmd == "train":
        wl = classes[a.cls]
        path = os.path.join(a.model_dir, os.path.basename(default_model_path(wl))) if a.model_dir \
            else default_model_path(wl)
        net, hist = train(wl, a.instances, seed_base=a.seed_base, ckpt=path, time_budget_s=a.budget)
        print(json.dumps({"model": path, "workload": wl.key(), "instances": hist[-1][0], "final_loss": hist[-1][1],
                          "first_loss": hist[0][1]}))
    else:
        paths = [os.path.join(a.model_dir, os.path.basename(default_model_path(wl))) for wl in classes] \
            if a.model_dir else None
        res = evaluate_batched(a.n, chunk=a.chunk, iters=[int(v) for v in a.iters.split(",")],
                               nets=load_nets(classes, paths))
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)
        print(json.dumps(res["summary"]))


if __name__ == "__main__":
    main()
