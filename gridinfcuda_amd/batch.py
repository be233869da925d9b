"""Batched BP (ginf_batch_*) and the seeded batched workloads.

A batch advances many independent single-level MRFs together: each
half-sweep of every instance of one label capacity is one grid-stride launch
over a work list, and the whole decode is captured once into a HIP graph and
replayed.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib as G

_declared = False


def _declare():
    global _declared
    if _declared:
        return
    L = G.lib()
    VP, I = ctypes.c_void_p, ctypes.c_int
    L.ginf_batch_create.argtypes = [I, ctypes.POINTER(G.BpDesc), ctypes.POINTER(VP), ctypes.POINTER(VP)]
    L.ginf_batch_run.argtypes = [VP, ctypes.POINTER(VP), ctypes.POINTER(VP), VP]
    L.ginf_batch_set_tiles.argtypes = [VP, ctypes.POINTER(I), I]
    L.ginf_batch_updates_per_run.argtypes = [VP, ctypes.POINTER(ctypes.c_double)]
    L.ginf_batch_dump_msgs.argtypes = [VP, I, VP]
    L.ginf_batch_destroy.argtypes = [VP]
    L.ginf_batch_set_seeded.argtypes = [VP, I]
    L.ginf_batch_seed.argtypes = [VP, I, VP, I, I, I, ctypes.c_float, VP]
    L.ginf_batch_set_compact_data.argtypes = [VP, I]
    L.ginf_batch_set_predictions.argtypes = [VP, ctypes.POINTER(VP), ctypes.POINTER(I), ctypes.POINTER(I), I,
                                             ctypes.POINTER(ctypes.c_float)]
    for n in ("ginf_batch_create", "ginf_batch_run", "ginf_batch_set_tiles", "ginf_batch_updates_per_run",
              "ginf_batch_dump_msgs", "ginf_batch_destroy", "ginf_batch_set_seeded", "ginf_batch_seed",
              "ginf_batch_set_compact_data", "ginf_batch_set_predictions"):
        getattr(L, n).restype = I
    _declared = True


LABEL_CLASSES = (16, 32, 64, 128)


class Batch:
    def __init__(self, descs, tables):
        _declare()
        self.n = len(descs)
        self.descs = descs
        self._tables = [np.ascontiguousarray(t) for t in tables]
        arr = (G.BpDesc * self.n)(*descs)
        tp = (ctypes.c_void_p * self.n)(*[t.ctypes.data for t in self._tables])
        self._h = ctypes.c_void_p()
        G.check(G.lib().ginf_batch_create(self.n, arr, tp, ctypes.byref(self._h)))
        self.tiles = ((256, 256, 256, 256), 4)

    def set_tiles(self, slots_per_item, blocks_per_cu: int) -> None:
        sp = (ctypes.c_int * 4)(*slots_per_item)
        G.check(G.lib().ginf_batch_set_tiles(self._h, sp, blocks_per_cu))
        self.tiles = (tuple(slots_per_item), blocks_per_cu)

    def run(self, data_ptrs, label_ptrs, stream: int = 0) -> None:
        dp = (ctypes.c_void_p * self.n)(*data_ptrs)
        lp = (ctypes.c_void_p * self.n)(*label_ptrs)
        G.check(G.lib().ginf_batch_run(self._h, dp, lp, ctypes.c_void_p(stream)))

    def set_compact_data(self, enable) -> None:
        """True/1: the narrowest data-cost storage that fits (8-bit, else
        16-bit, else T) per instance; 2: at most 16-bit; False/0: T."""
        G.check(G.lib().ginf_batch_set_compact_data(self._h, int(enable)))

    def set_seeded(self, seeded: bool) -> None:
        G.check(G.lib().ginf_batch_set_seeded(self._h, 1 if seeded else 0))

    def seed(self, i: int, pred_ptr: int, pw: int, ph: int, shift: int, scale: float, stream: int = 0) -> None:
        G.check(G.lib().ginf_batch_seed(self._h, i, ctypes.c_void_p(pred_ptr), pw, ph, shift, scale,
                                        ctypes.c_void_p(stream)))

    def set_predictions(self, preds, shift: int = 2) -> None:
        """preds: one (pointer, pw, ph, scale) per instance, or None to switch
        off: the first half-sweep of every run reads instance i's incoming
        messages from its prediction (ginf_batch_set_predictions)."""
        if preds is None:
            G.check(G.lib().ginf_batch_set_predictions(self._h, None, None, None, 0, None))
            return
        n = self.n
        ptr = (ctypes.c_void_p * n)(*[p[0] for p in preds])
        pw = (ctypes.c_int * n)(*[p[1] for p in preds])
        ph = (ctypes.c_int * n)(*[p[2] for p in preds])
        sc = (ctypes.c_float * n)(*[p[3] for p in preds])
        G.check(G.lib().ginf_batch_set_predictions(self._h, ptr, pw, ph, shift, sc))

    def updates_per_run(self) -> float:
        u = ctypes.c_double()
        G.check(G.lib().ginf_batch_updates_per_run(self._h, ctypes.byref(u)))
        return u.value

    def dump_msgs(self, i: int) -> np.ndarray:
        d = self.descs[i]
        out = np.zeros((d.num_labels, 4, d.height, d.width), np.float32 if d.dtype == G.F32 else np.int32)
        G.check(G.lib().ginf_batch_dump_msgs(self._h, i, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def close(self) -> None:
        if self._h:
            G.lib().ginf_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014): the stream the instance
    parameters of the batched workloads are drawn from."""

    M = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = seed & self.M

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.M
        return z ^ (z >> 31)

    def unit(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))


def table_potts(L, w, dtype=np.float32):
    a = np.arange(L)
    return (w * (a[:, None] != a[None, :])).astype(dtype)


def table_trunc_linear(L, scale, d, dtype=np.float32):
    a = np.arange(L)
    return np.minimum(scale * np.abs(a[:, None] - a[None, :]), d).astype(dtype)


def mixed_instances(n=64, base_seed=3000, wmin=160, wmax=1280, hmin=120, hmax=720):
    """The batched workload: n instances, widths log-uniform in [wmin, wmax],
    heights log-uniform in [hmin, hmax], label counts from {16,32,64,128};
    Potts (w=15) on even indices, truncated linear lambda=8 tau=4 on odd ones.
    Instance i's costs come from the stereo generator with seed base_seed + i
    (16 rectangles, disparities 0..L-1, truncation 40)."""
    rng = _SplitMix64(base_seed)
    out = []
    for i in range(n):
        W = int(round(math.exp(math.log(wmin) + rng.unit() * (math.log(wmax) - math.log(wmin)))))
        H = int(round(math.exp(math.log(hmin) + rng.unit() * (math.log(hmax) - math.log(hmin)))))
        L = LABEL_CLASSES[rng.next() % 4]
        potts = i % 2 == 0
        out.append(dict(index=i, seed=base_seed + i, W=W, H=H, L=L, model=G.SM_POTTS if potts else G.SM_TRUNC_LINEAR,
                        table=table_potts(L, 15) if potts else table_trunc_linear(L, 8, 32)))
    return out


def instance_costs(inst, dtype=np.float32):
    key = "cost_f32" if np.dtype(dtype) == np.float32 else "cost_i32"
    return G.gen_stereo(inst["seed"], inst["W"], inst["H"], inst["L"], 16, inst["L"] - 1, 40, want=(key,))[key]


def make_batch(instances, iters=20, norm=G.NORM_MEAN, dtype=np.float32) -> Batch:
    descs, tables = [], []
    for inst in instances:
        descs.append(G.make_desc(inst["W"], inst["H"], inst["L"], G.dtype_code(dtype), inst["model"], 1, iters, norm))
        tables.append(np.ascontiguousarray(inst["table"], dtype=dtype))
    return Batch(descs, tables)
