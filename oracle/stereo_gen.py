"""ORACLE -- numpy restatement of the seeded synthetic stereo generator
(project-owned; the product's C++ version is ginf_gen_stereo).  Used only by
tests to pin the product generator byte for byte.

Same specification, written differently (vectorised filter, literal
paint-in-increasing-disparity warp):
  SplitMix64 stream from `seed`; uniform int in [lo, hi] = lo + next % (hi-lo+1);
  noise byte = next >> 56.
  left  = 3x3 box filter (clamp-to-edge) of W*H noise bytes, (sum + 4) // 9
  disp  = min(2, dmax) background, then num_rects rectangles
          (x0, y0, w in [max(1,W//16), max(.., W//4)], h likewise, d in [0, dmax])
  right(x - d, y) = left(x, y), painted for d = 0, 1, ..., dmax; holes = noise
  cost(l, y, x) = min(|left(x,y) - right(x-l,y)|, trunc), trunc where x < l
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & M64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def uniform(self, lo: int, hi: int) -> int:
        return lo + self.next() % (hi - lo + 1)

    def byte(self) -> int:
        return self.next() >> 56


def gen_stereo(seed, W, H, L, num_rects, dmax, trunc):
    rng = SplitMix64(seed)
    raw = np.array([rng.byte() for _ in range(W * H)], np.int64).reshape(H, W)
    pad = np.pad(raw, 1, mode="edge")
    acc = sum(pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    left = ((acc + 4) // 9).astype(np.uint8)
    disp = np.full((H, W), min(2, dmax), np.int32)
    wmin = max(1, W // 16)
    wmax = max(wmin, W // 4)
    hmin = max(1, H // 16)
    hmax = max(hmin, H // 4)
    for _ in range(num_rects):
        x0 = rng.uniform(0, W - 1)
        y0 = rng.uniform(0, H - 1)
        rw = rng.uniform(wmin, wmax)
        rh = rng.uniform(hmin, hmax)
        d = rng.uniform(0, dmax)
        disp[y0:y0 + rh, x0:x0 + rw] = d
    right = np.zeros((H, W), np.uint8)
    filled = np.zeros((H, W), bool)
    xs = np.arange(W)
    for d in range(dmax + 1):
        ys, xsel = np.nonzero((disp == d) & (xs[None, :] >= d))
        right[ys, xsel - d] = left[ys, xsel]
        filled[ys, xsel - d] = True
    for y in range(H):
        for x in range(W):
            if not filled[y, x]:
                right[y, x] = rng.byte()
    cost = np.full((L, H, W), trunc, np.int32)
    li = left.astype(np.int32)
    ri = right.astype(np.int32)
    for l in range(L):
        if l < W:
            cost[l, :, l:] = np.minimum(np.abs(li[:, l:] - ri[:, :W - l]), trunc)
    return {"left": left, "right": right, "disp": disp, "cost_i32": cost}
