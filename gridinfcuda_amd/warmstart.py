"""Learned warm start for hierarchical BP (project-owned; no reference
counterpart).

A 6-layer 3x3 convolutional network (BF16 weights and activations, FP32
accumulation, 32 hidden channels, ReLU) maps the data-cost volume min-pooled
to 1/4 resolution (L input channels) to the four directional message volumes
at 1/4 resolution (4L output channels).  Its prediction seeds a fine level of
the BP pyramid (outbox-frame upsample, border outboxes kept at zero), which
then runs a few red-black half-sweeps; plain hierarchical BP stays the
fallback and the energy reference.

Every layer runs on the library's own HIP kernels (ginf_conv3x3_* on MFMA);
PyTorch only provides device memory and the stream.  Training targets are the
converged level-0 messages of plain HBP on seeded synthetic stereo
instances, averaged over 4x4 blocks.
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes
import json
import os
import time

import numpy as np

from . import _lib as G

PKG = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(PKG, "models")

POOL = 4          # input / output resolution factor
IN_SCALE = 1.0 / 40.0   # cost scale (generator truncation 40)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _ck(rc):
    G.check(rc)


class ConvNet:
    """Weights [Cout][3][3][Cin] (FP32 master + BF16 copy), biases FP32, Adam state."""

    def __init__(self, labels: int, hidden: int = 32, layers: int = 6, seed: int = 0):
        import torch
        self.torch = torch
        G.lib()
        self.L = labels
        self.hidden = hidden
        self.chans = [labels] + [hidden] * (layers - 1) + [4 * labels]
        rng = np.random.default_rng(seed)
        self.w, self.b, self.wb, self.wf = [], [], [], []
        self.m1, self.m2 = [], []
        for k in range(layers):
            ci, co = self.chans[k], self.chans[k + 1]
            std = np.sqrt(2.0 / (9 * ci)) * (0.5 if k == layers - 1 else 1.0)
            w = torch.from_numpy((rng.standard_normal((co, 3, 3, ci)) * std).astype(np.float32)).cuda()
            b = torch.zeros(co, dtype=torch.float32, device="cuda")
            self.w.append(w)
            self.b.append(b)
            self.wb.append(torch.empty(w.shape, dtype=torch.int16, device="cuda"))
            self.wf.append(torch.empty((ci, 3, 3, co), dtype=torch.int16, device="cuda"))
            self.m1.append([torch.zeros_like(w), torch.zeros_like(b)])
            self.m2.append([torch.zeros_like(w), torch.zeros_like(b)])
        self.step_count = 0
        self.refresh()
        self._bufs = {}

    @property
    def layers(self):
        return len(self.w)

    def stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def refresh(self):
        """BF16 weight copies and the flipped kernels used by the input gradient."""
        for k in range(self.layers):
            _ck(G.lib().ginf_to_bf16(_vp(self.w[k]), self.w[k].numel(), _vp(self.wb[k]), self.stream()))
            _ck(G.lib().ginf_conv3x3_flip_weights(_vp(self.wb[k]), self.chans[k + 1], self.chans[k], _vp(self.wf[k]),
                                                 self.stream()))

    def params(self) -> int:
        return int(sum(w.numel() + b.numel() for w, b in zip(self.w, self.b)))

    def _buf(self, key, shape, dtype):
        t = self._bufs.get(key)
        if t is None or tuple(t.shape) != tuple(shape) or t.dtype != dtype:
            t = self.torch.empty(shape, dtype=dtype, device="cuda")
            self._bufs[key] = t
        return t

    def forward(self, x_bf16, N: int, H: int, W: int, out=None):
        """x_bf16: int16 tensor (N,H,W,L) holding bf16.  Returns FP32 (N,H,W,4L);
        keeps the layer inputs for backward."""
        torch = self.torch
        self.acts = [x_bf16]
        h = x_bf16
        for k in range(self.layers):
            ci, co = self.chans[k], self.chans[k + 1]
            last = k == self.layers - 1
            if last:
                y = out if out is not None else self._buf(("out",), (N, H, W, co), torch.float32)
                _ck(G.lib().ginf_conv3x3_fwd(_vp(h), _vp(self.wb[k]), _vp(self.b[k]), N, H, W, ci, co, 0, None,
                                            _vp(y), self.stream()))
                self.out = y
            else:
                y = self._buf(("act", k), (N, H, W, co), torch.int16)
                _ck(G.lib().ginf_conv3x3_fwd(_vp(h), _vp(self.wb[k]), _vp(self.b[k]), N, H, W, ci, co, 1, _vp(y),
                                            None, self.stream()))
                self.acts.append(y)
                h = y
        self.shape = (N, H, W)
        return self.out

    def infer(self, x_bf16, N: int, H: int, W: int, out):
        """Inference forward in one library call (ginf_convnet_fwd): the
        same kernels and results as forward(), without keeping activations."""
        torch = self.torch
        hid = max(self.chans[1:-1]) if self.layers > 1 else 1
        a0 = self._buf(("inf", 0), (N, H, W, hid), torch.int16)
        a1 = self._buf(("inf", 1), (N, H, W, hid), torch.int16)
        key = tuple(t.data_ptr() for t in self.wb + self.b)
        if getattr(self, "_ptrs_key", None) != key:
            n = self.layers
            self._c_chans = (ctypes.c_int * (n + 1))(*self.chans)
            self._c_w = (ctypes.c_void_p * n)(*[t.data_ptr() for t in self.wb])
            self._c_b = (ctypes.c_void_p * n)(*[t.data_ptr() for t in self.b])
            self._ptrs_key = key
        _ck(G.lib().ginf_convnet_fwd(self.layers, self._c_chans, self._c_w, self._c_b, _vp(x_bf16), N, H, W, _vp(a0),
                                    _vp(a1), _vp(out), self.stream()))
        return out

    def backward(self, dout):
        """dout: FP32 gradient of the loss w.r.t. the output.  Fills self.gw/self.gb."""
        torch = self.torch
        N, H, W = self.shape
        self.gw = [None] * self.layers
        self.gb = [None] * self.layers
        dy = dout
        for k in range(self.layers - 1, -1, -1):
            ci, co = self.chans[k], self.chans[k + 1]
            last = k == self.layers - 1
            dz = self._buf(("dz", k), (N, H, W, co), torch.int16)
            yk = None if last else self.acts[k + 1]
            _ck(G.lib().ginf_relu_bwd(_vp(dy), _vp(yk) if yk is not None else None, dy.numel(), _vp(dz),
                                     self.stream()))
            gw = self._buf(("gw", k), tuple(self.w[k].shape), torch.float32)
            gb = self._buf(("gb", k), (co,), torch.float32)
            _ck(G.lib().ginf_conv3x3_bwd_weight(_vp(self.acts[k]), _vp(dz), N, H, W, ci, co, _vp(gw), _vp(gb),
                                               self.stream()))
            self.gw[k], self.gb[k] = gw, gb
            if k > 0:
                dx = self._buf(("dx", k), (N, H, W, ci), torch.float32)
                _ck(G.lib().ginf_conv3x3_fwd(_vp(dz), _vp(self.wf[k]), None, N, H, W, co, ci, 0, None, _vp(dx),
                                            self.stream()))
                dy = dx

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step_count += 1
        for k in range(self.layers):
            _ck(G.lib().ginf_adam(_vp(self.w[k]), _vp(self.gw[k]), _vp(self.m1[k][0]), _vp(self.m2[k][0]),
                                 _vp(self.wb[k]), self.w[k].numel(), lr, b1, b2, eps, self.step_count, wd,
                                 self.stream()))
            _ck(G.lib().ginf_adam(_vp(self.b[k]), _vp(self.gb[k]), _vp(self.m1[k][1]), _vp(self.m2[k][1]), None,
                                 self.b[k].numel(), lr, b1, b2, eps, self.step_count, 0.0, self.stream()))
        for k in range(self.layers):
            _ck(G.lib().ginf_conv3x3_flip_weights(_vp(self.wb[k]), self.chans[k + 1], self.chans[k], _vp(self.wf[k]),
                                                 self.stream()))

    def mse(self, target):
        """Loss and its gradient (written into a buffer) for the last forward."""
        torch = self.torch
        grad = self._buf(("gout",), tuple(self.out.shape), torch.float32)
        parts = self._buf(("parts",), (1024,), torch.float32)
        nparts = ctypes.c_int()
        _ck(G.lib().ginf_mse(_vp(self.out), _vp(target), self.out.numel(), _vp(grad), _vp(parts),
                            ctypes.byref(nparts), self.stream()))
        loss = parts[:nparts.value].sum() / self.out.numel()
        return loss, grad

    # ---- checkpoints ------------------------------------------------------
    def state(self) -> dict:
        d = {"L": self.L, "hidden": self.hidden, "layers": self.layers, "step": self.step_count}
        arrs = {}
        for k in range(self.layers):
            arrs[f"w{k}"] = self.w[k].cpu().numpy()
            arrs[f"b{k}"] = self.b[k].cpu().numpy()
            arrs[f"m1w{k}"] = self.m1[k][0].cpu().numpy()
            arrs[f"m1b{k}"] = self.m1[k][1].cpu().numpy()
            arrs[f"m2w{k}"] = self.m2[k][0].cpu().numpy()
            arrs[f"m2b{k}"] = self.m2[k][1].cpu().numpy()
        return d, arrs

    def save(self, path: str, extra: dict | None = None) -> None:
        meta, arrs = self.state()
        if extra:
            meta.update(extra)
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = path + ".tmp.npz"
        np.savez(tmp, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **arrs)
        os.replace(tmp, path)

    @classmethod
    def load(cls, path: str) -> "ConvNet":
        import torch
        z = np.load(path, allow_pickle=False)
        meta = json.loads(bytes(z["meta"]).decode())
        net = cls(meta["L"], meta["hidden"], meta["layers"])
        for k in range(net.layers):
            net.w[k].copy_(torch.from_numpy(z[f"w{k}"]))
            net.b[k].copy_(torch.from_numpy(z[f"b{k}"]))
            if f"m1w{k}" in z:
                net.m1[k][0].copy_(torch.from_numpy(z[f"m1w{k}"]))
                net.m1[k][1].copy_(torch.from_numpy(z[f"m1b{k}"]))
                net.m2[k][0].copy_(torch.from_numpy(z[f"m2w{k}"]))
                net.m2[k][1].copy_(torch.from_numpy(z[f"m2b{k}"]))
        net.step_count = meta.get("step", 0)
        net.meta = meta
        net.refresh()
        return net


# ---------------------------------------------------------------- instances

class Workload:
    """A warm-start workload: grid size, labels, smoothness (truncated linear
    lambda * min(|a-b|, tau) or Potts), generator parameters, BP schedule."""

    def __init__(self, W=640, H=480, L=64, model=G.SM_TRUNC_LINEAR, lam=10, tau=3, rects=24, trunc=40,
                 levels=5, iters=10):
        self.W, self.H, self.L = W, H, L
        self.model, self.lam, self.tau = model, lam, tau
        self.rects, self.trunc = rects, trunc
        self.levels, self.iters = levels, iters
        a = np.arange(L)
        if model == G.SM_POTTS:
            self.table = (lam * (a[:, None] != a[None, :])).astype(np.float32)
        else:
            self.table = np.minimum(lam * np.abs(a[:, None] - a[None, :]), lam * tau).astype(np.float32)
        self.msg_scale = 1.0 / float(max(1.0, lam * (tau if model != G.SM_POTTS else 1)))
        self.w4, self.h4 = (W + POOL - 1) // POOL, (H + POOL - 1) // POOL

    def key(self):
        return f"{self.W}x{self.H}x{self.L}-{'potts' if self.model == G.SM_POTTS else 'tl'}{self.lam}-{self.tau}"


def gen_images(seeds, wl: Workload, threads: int = 16):
    """Left/right uint8 images of seeded instances (host generator, threaded)."""
    def one(seed):
        g = G.gen_stereo(seed, wl.W, wl.H, wl.L, wl.rects, wl.L - 1, wl.trunc, want=("left", "right"))
        return g["left"], g["right"]
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, seeds))


class Pipeline:
    """Device-side instance preparation for one workload: costs from images,
    plain HBP (the energy reference), targets and network inputs."""

    def __init__(self, wl: Workload):
        import torch
        self.torch = torch
        self.wl = wl
        self.plan = G.Plan(G.make_desc(wl.W, wl.H, wl.L, G.F32, wl.model, wl.levels, wl.iters, G.NORM_MEAN),
                           wl.table)
        self.costs = torch.empty((wl.L, wl.H, wl.W), dtype=torch.float32, device="cuda")
        self.labels = torch.zeros((wl.H, wl.W), dtype=torch.int32, device="cuda")
        self.left = torch.empty((wl.H, wl.W), dtype=torch.uint8, device="cuda")
        self.right = torch.empty((wl.H, wl.W), dtype=torch.uint8, device="cuda")
        self.table_dev = torch.from_numpy(wl.table).cuda()

    def stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def load_costs(self, left: np.ndarray, right: np.ndarray):
        wl = self.wl
        self.left.copy_(self.torch.from_numpy(left))
        self.right.copy_(self.torch.from_numpy(right))
        _ck(G.lib().ginf_stereo_costs(_vp(self.left), _vp(self.right), wl.W, wl.H, wl.L, wl.trunc, _vp(self.costs),
                                     ctypes.c_void_p(self.stream())))

    def net_input(self, out_slot):
        wl = self.wl
        _ck(G.lib().ginf_minpool_costs(_vp(self.costs), G.F32, wl.W, wl.H, wl.L, POOL, IN_SCALE, _vp(out_slot),
                                      ctypes.c_void_p(self.stream())))

    def plain(self):
        self.plan.run(self.costs.data_ptr(), self.labels.data_ptr(), self.stream())

    def target(self, out_slot, level=0):
        """Outboxes of `level` after the plain run, averaged over the blocks
        of level nodes that one 1/4-resolution cell covers (4x4 at level 0,
        2x2 at level 1, one node at level 2)."""
        self.plan.block_mean_msgs(level, POOL >> level, self.wl.msg_scale, out_slot.data_ptr(), self.stream())

    def energy(self) -> float:
        wl = self.wl
        return G.energy_f64(wl.W, wl.H, wl.L, G.F32, self.costs.data_ptr(), self.table_dev.data_ptr(),
                            self.labels.data_ptr(), self.stream())


# ---------------------------------------------------------------- training

def train(wl: Workload, n_instances=20000, batch=16, lr=2e-3, reuse=2, seed_base=1_000_000, ckpt=None,
          log=print, resume=None, time_budget_s=None, target_level=0, hidden=32, layers=6):
    """Online training: each batch of fresh seeded instances is labelled by
    plain HBP on the GPU and used for `reuse` Adam steps.  The targets are
    plain HBP's outboxes of `target_level` (0: the converged finest level,
    averaged over 4x4 blocks; 2: the 1/4-resolution level at the end of its
    half-sweeps, the state that plain HBP hands to level 1)."""
    import torch
    net = ConvNet.load(resume) if resume else ConvNet(wl.L, hidden, layers)
    if resume:
        target_level = net.meta.get("target_level", 0)
    pipe = Pipeline(wl)
    xin = torch.empty((batch, wl.h4, wl.w4, wl.L), dtype=torch.int16, device="cuda")
    tgt = torch.empty((batch, wl.h4, wl.w4, 4 * wl.L), dtype=torch.float32, device="cuda")
    done = net.meta.get("instances", 0) if resume else 0
    t0 = time.time()
    history = []
    pool = cf.ThreadPoolExecutor(max_workers=1)
    seeds = list(range(seed_base + done, seed_base + n_instances))
    nxt = pool.submit(gen_images, seeds[:batch], wl)
    pos = 0
    while pos < len(seeds):
        imgs = nxt.result()
        cur = seeds[pos:pos + batch]
        pos += len(cur)
        if pos < len(seeds):
            nxt = pool.submit(gen_images, seeds[pos:pos + batch], wl)
        nb = len(cur)
        for k, (lft, rgt) in enumerate(imgs):
            pipe.load_costs(lft, rgt)
            pipe.net_input(xin[k])
            pipe.plain()
            pipe.target(tgt[k], target_level)
        nbatches = max(1, len(seeds) // batch)
        warm = min(20, nbatches // 5 + 1)
        step_lr = lr * min(1.0, (pos / batch) / warm) * (0.1 ** (pos / max(1, len(seeds))))
        for _ in range(reuse):
            net.forward(xin[:nb], nb, wl.h4, wl.w4)
            loss, g = net.mse(tgt[:nb])
            net.backward(g)
            net.adam(lr=step_lr)
        done += nb
        lv = float(loss)
        history.append((done, lv))
        if len(history) % 25 == 1 or pos >= len(seeds):
            log(f"  instances {done:6d}  loss {lv:.5f}  lr {step_lr:.2e}  {time.time() - t0:.1f}s")
        if ckpt and len(history) % 200 == 0:
            net.save(ckpt, {"instances": done, "workload": wl.key(), "target_level": target_level})
        if time_budget_s and time.time() - t0 > time_budget_s:
            log(f"  time budget reached after {done} instances")
            break
    if ckpt:
        net.save(ckpt, {"instances": done, "workload": wl.key(), "loss": history[-1][1],
                        "target_level": target_level})
    return net, history


# ---------------------------------------------------------------- inference

class WarmStart:
    """Network inference + seeded BP on one workload.  `seed_level` is the
    pyramid level seeded from the 1/4-resolution prediction (0: finest, with
    a 4x outbox upsample; 1: half resolution, 2x)."""

    def __init__(self, wl: Workload, net: ConvNet | None):
        import torch
        self.torch = torch
        self.wl = wl
        self.net = net
        self.plan = G.Plan(G.make_desc(wl.W, wl.H, wl.L, G.F32, wl.model, max(2, wl.levels), wl.iters, G.NORM_MEAN),
                           wl.table)
        self.xin = torch.empty((1, wl.h4, wl.w4, wl.L), dtype=torch.int16, device="cuda")
        self.pred = torch.zeros((1, wl.h4, wl.w4, 4 * wl.L), dtype=torch.float32, device="cuda")

    def stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def run(self, costs, labels, schedule, seed_level=1):
        """schedule: half-sweeps per level for levels seed_level..0 (index =
        level).  seed_level < 0: plain HBP on all levels, no network."""
        wl = self.wl
        if seed_level < 0:
            self.plan.run_schedule(costs.data_ptr(), labels.data_ptr(), len(schedule) - 1, schedule, False,
                                   self.stream())
            return
        if self.net is None:
            # zero messages at the seeded level: an unseeded start there (the
            # same result as seeding zeros, without storing or reading them)
            self.plan.run_schedule(costs.data_ptr(), labels.data_ptr(), seed_level, schedule, False, self.stream())
            return
        # One pass over the costs builds the level data and the network input
        # (ginf_bp_plan_prepare); the network runs in one launch; the first
        # half-sweep of the seeded level reads the prediction itself (same
        # result as ginf_bp_plan_seed followed by a seeded schedule).
        self.plan.prepare(costs.data_ptr(), seed_level, self.xin.data_ptr(), IN_SCALE, self.stream())
        self.net.infer(self.xin, 1, wl.h4, wl.w4, self.pred)
        self.plan.run_predicted(costs.data_ptr(), labels.data_ptr(), seed_level, schedule, self.pred.data_ptr(), wl.w4,
                                wl.h4, 2 - seed_level, wl.msg_scale, self.stream(), prepared=True)


def default_model_path(wl: Workload) -> str:
    return os.path.join(MODELS, f"warmstart_{wl.key()}.npz")


# ---------------------------------------------------------------- evaluation

def _timed(torch, fn, reps=5):
    st = torch.cuda.current_stream()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        fn()
        b.record(st)
        b.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts))


# (seed level, half-sweeps per level from the seeded level down to level 0);
# seed level -1 is plain HBP from zero messages on all 5 levels (a baseline
# with fewer iterations than the reference's 10 per level).
SCHEDULES = [(0, (4,)), (0, (6,)), (0, (8,)), (0, (10,)), (0, (11,)), (0, (12,)), (0, (13,)), (0, (14,)),
             (1, (2, 4)), (1, (4, 4)), (1, (4, 6)), (1, (6, 6)), (1, (6, 8)), (1, (0, 10)), (1, (2, 10)), (1, (3, 10)),
             (1, (4, 10)), (1, (6, 10)), (1, (8, 10)), (1, (10, 10)), (1, (4, 8)), (1, (6, 9)),
             (2, (4, 4, 6)), (2, (6, 6, 6)), (2, (6, 6, 8)), (2, (0, 8, 10)), (2, (0, 10, 8)), (2, (0, 10, 10)),
             (2, (1, 10, 10)), (2, (2, 8, 10)), (2, (2, 10, 10)), (2, (4, 10, 10)), (2, (6, 10, 10)), (2, (10, 10, 10)),
             (-1, (4, 4, 4, 4, 4)), (-1, (6, 6, 6, 6, 6)), (-1, (10, 10, 10, 6, 6)), (-1, (10, 10, 10, 10, 8))]


def evaluate(wl: Workload, net: ConvNet | None, seeds, schedules=SCHEDULES, reps=5, log=print):
    """Energy and GPU time of plain HBP (the energy reference) against seeded schedules,
    per instance.  A schedule (level, iters) seeds `level` from the network
    (or from zeros when net is None) and runs iters[t] half-sweeps at level
    level-t ... 0."""
    import torch
    pipe = Pipeline(wl)
    ws = WarmStart(wl, net)
    lab = torch.zeros((wl.H, wl.W), dtype=torch.int32, device="cuda")
    imgs = gen_images(list(seeds), wl)
    out = []
    for seed, (lft, rgt) in zip(seeds, imgs):
        pipe.load_costs(lft, rgt)
        t_plain = _timed(torch, pipe.plain, reps)
        e_plain = pipe.energy()
        rec = {"seed": seed, "plain_ms": t_plain, "plain_energy": e_plain, "warm": []}
        for level, iters in schedules:
            sched = list(reversed(iters))

            def go():
                ws.run(pipe.costs, lab, sched, seed_level=level)
            t = _timed(torch, go, reps)
            e = G.energy_f64(wl.W, wl.H, wl.L, G.F32, pipe.costs.data_ptr(), pipe.table_dev.data_ptr(),
                             lab.data_ptr(), torch.cuda.current_stream().cuda_stream)
            rec["warm"].append({"level": level, "iters": list(iters), "ms": t, "energy": e,
                                "rel_excess": e / e_plain - 1.0})
        out.append(rec)
        best = min((w for w in rec["warm"] if w["rel_excess"] <= 1e-3), key=lambda w: w["ms"], default=None)
        log(f"  seed {seed}: plain {t_plain:.3f} ms E={e_plain:.0f}; fastest within 0.1%: "
            + (f"{best['level']}/{best['iters']} {best['ms']:.3f} ms ({best['ms'] / t_plain:.2f}x)" if best else "none"))
    return out


def summarize(records, schedules=SCHEDULES):
    """Per schedule: mean time ratio to plain, fraction of instances within 0.1 %."""
    rows = []
    for k, (level, iters) in enumerate(schedules):
        ratios = [r["warm"][k]["ms"] / r["plain_ms"] for r in records]
        ok = [r["warm"][k]["rel_excess"] <= 1e-3 for r in records]
        exc = [r["warm"][k]["rel_excess"] for r in records]
        rows.append({"level": level, "iters": list(iters), "time_ratio": float(np.mean(ratios)),
                     "within_0.1pct": float(np.mean(ok)), "mean_rel_excess": float(np.mean(exc))})
    return rows


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="train / evaluate the learned warm start")
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--W", type=int, default=640)
    ap.add_argument("--H", type=int, default=480)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--potts", action="store_true")
    ap.add_argument("--lam", type=float, default=10)
    ap.add_argument("--tau", type=float, default=3)
    ap.add_argument("--instances", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--reuse", type=int, default=2)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--model", default=None)
    ap.add_argument("--resume", default=None)
    ap.add_argument("--budget", type=float, default=None, help="training time budget in seconds")
    ap.add_argument("--target-level", type=int, default=0, choices=[0, 1, 2],
                    help="train: pyramid level whose plain-HBP outboxes are the targets")
    ap.add_argument("--hidden", type=int, default=32, help="train: hidden channels (a multiple of 32)")
    ap.add_argument("--layers", type=int, default=6, help="train: convolution layers")
    ap.add_argument("--eval-seeds", type=int, default=8)
    ap.add_argument("--out", default=None, help="eval: write JSON records here")
    a = ap.parse_args(argv)
    wl = Workload(a.W, a.H, a.L, G.SM_POTTS if a.potts else G.SM_TRUNC_LINEAR, a.lam, a.tau)
    path = a.model or default_model_path(wl)
    if a.cmd == "train":
        net, hist = train(wl, a.instances, a.batch, a.lr, a.reuse, ckpt=path, resume=a.resume, time_budget_s=a.budget,
                          target_level=a.target_level, hidden=a.hidden, layers=a.layers)
        print(json.dumps({"model": path, "params": net.params(), "instances": hist[-1][0], "final_loss": hist[-1][1],
                          "first_loss": hist[0][1]}))
    else:
        net = ConvNet.load(path)
        seeds = list(range(2_000_000, 2_000_000 + a.eval_seeds))
        rec_net = evaluate(wl, net, seeds)
        rec_zero = evaluate(wl, None, seeds)
        res = {"workload": wl.key(), "model": path, "with_net": summarize(rec_net), "zero_seed": summarize(rec_zero),
               "records": rec_net}
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)
        for name in ("with_net", "zero_seed"):
            print(name)
            for r in res[name]:
                print(f"  seed level {r['level']} iters {r['iters']}: time {r['time_ratio']:.2f}x plain, "
                      f"within 0.1%: {r['within_0.1pct']:.2f}, mean excess {r['mean_rel_excess']:.2e}")


if __name__ == "__main__":
    main()
