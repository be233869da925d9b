"""Stereo front-end: the reference's example pipeline (examples/stereo/
stereo.cc:15-51) on this library.

Load a rectified image pair as 8-bit grey, build the 16-label data costs
min(|L(x,y) - R(x-k,y)|, 20) for x >= 15 (other columns keep the zero the
Grid starts with), truncated-linear smoothness min(10|a-b|, 20), decodeBp
with 30 half-sweeps on the GPU, disparity image = label * (256 / 16).
Optionally refine with ICM (the reference's adjacent decoder).

The reference loads with OpenCV (CV_LOAD_IMAGE_GRAYSCALE), which is absent
here; RGB images are converted with the ITU-R BT.601 weights in 14-bit fixed
point, (4899 R + 9617 G + 1868 B + 8192) >> 14, which is OpenCV's
cvtColor(RGB2GRAY) rule.

    python -m gridinfcuda_amd.stereo left.png right.png -o disparity.png
"""
from __future__ import annotations

import argparse

import numpy as np

from . import _lib as G

LABELS = 16
TRUNC = 20
LAMBDA, D = 10, 20
ITERS = 30


def load_gray(path: str) -> np.ndarray:
    from PIL import Image
    im = Image.open(path)
    a = np.asarray(im)
    if a.ndim == 2:
        return a.astype(np.uint8)
    rgb = a[..., :3].astype(np.uint32)
    g = (4899 * rgb[..., 0] + 9617 * rgb[..., 1] + 1868 * rgb[..., 2] + 8192) >> 14
    return g.astype(np.uint8)


def example_costs(left: np.ndarray, right: np.ndarray, labels: int = LABELS, trunc: int = TRUNC,
                  dtype=np.int32) -> np.ndarray:
    """Data costs [L][H][W] of examples/stereo/stereo.cc:28-34."""
    H, W = left.shape
    lft, rgt = left.astype(np.int32), right.astype(np.int32)
    c = np.zeros((labels, H, W), dtype)
    x0 = labels - 1
    for k in range(labels):
        c[k, :, x0:] = np.minimum(np.abs(lft[:, x0:] - rgt[:, x0 - k:W - k]), trunc)
    return c


def smoothness_table(labels: int = LABELS, lam: int = LAMBDA, d: int = D, dtype=np.int32) -> np.ndarray:
    """useSmoothnessTruncLinear(lam, d): V(a,b) = min(lam*|a-b|, d)."""
    a = np.arange(labels)
    return np.minimum(lam * np.abs(a[:, None] - a[None, :]), d).astype(dtype)


def disparity(left: np.ndarray, right: np.ndarray, iters: int = ITERS, levels: int = 1, icm_iters: int = 0,
              icm_mode: int = G.ICM_ASYNC) -> np.ndarray:
    """Labels [H][W] of the example pipeline (Grid<int>, BP on the GPU)."""
    costs = example_costs(left, right)
    table = smoothness_table()
    lab = G.hbp_host(costs, table, G.SM_TRUNC_LINEAR, levels, iters)
    if icm_iters:
        lab = G.icm_host(costs, table, lab, icm_iters, icm_mode)
    return lab


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="stereo disparity with grid BP on the GPU")
    ap.add_argument("left")
    ap.add_argument("right")
    ap.add_argument("-o", "--out", default="disparity.png")
    ap.add_argument("--iters", type=int, default=ITERS)
    ap.add_argument("--levels", type=int, default=1)
    ap.add_argument("--icm", type=int, default=0, help="ICM refinement iterations after BP")
    a = ap.parse_args(argv)
    left, right = load_gray(a.left), load_gray(a.right)
    if left.shape != right.shape:
        print("Error in loading stereo image pair")
        return 1
    lab = disparity(left, right, a.iters, a.levels, a.icm)
    from PIL import Image
    Image.fromarray((lab * (256 // LABELS)).astype(np.uint8)).save(a.out)
    import torch
    costs = torch.from_numpy(example_costs(left, right)).cuda()
    table = torch.from_numpy(smoothness_table()).cuda()
    labd = torch.from_numpy(lab).cuda()
    e = G.energy_f64(left.shape[1], left.shape[0], LABELS, G.I32, costs.data_ptr(), table.data_ptr(),
                     labd.data_ptr(), torch.cuda.current_stream().cuda_stream)
    print(f"{left.shape[1]}x{left.shape[0]}: energy {e:.0f} -> {a.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
