"""Helpers for the GPU tests: run the HIP path through the C ABI with
device buffers from torch (torch is only the allocator and stream)."""
from __future__ import annotations

import sys

import numpy as np

from gridinfcuda_amd import _lib as G


def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def run_plan(data: np.ndarray, table: np.ndarray, model: int, levels: int, iters: int, norm: int = 0,
             lanes: int = 0, labels_init: np.ndarray | None = None, dump_level: int | None = None,
             runs: int = 1, kernel: int = 0, compact: bool | None = None, pairs: bool | None = None):
    """HBP on the GPU via ginf_bp_plan_*; returns (labels, msgs-or-None, plan).
    compact: None keeps the plan's default (16-bit data costs when they fit);
    pairs: None keeps the plan default for fused pairs of half-sweeps."""
    torch = torch_cuda()
    data = np.ascontiguousarray(data)
    L, H, W = data.shape
    desc = G.make_desc(W, H, L, G.dtype_code(data.dtype), model, levels, iters, norm, lanes, kernel)
    plan = G.Plan(desc, np.ascontiguousarray(table, dtype=data.dtype))
    if compact is not None:
        plan.set_compact_data(compact)
    if pairs is not None:
        plan.set_pairs(pairs)
    dd = torch.from_numpy(data).cuda()
    init = np.zeros((H, W), np.int32) if labels_init is None else labels_init.astype(np.int32)
    out = None
    for _ in range(runs):
        ld = torch.from_numpy(np.ascontiguousarray(init)).cuda()
        plan.run(dd.data_ptr(), ld.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        lab = ld.cpu().numpy()
        if out is not None:
            np.testing.assert_array_equal(lab, out, err_msg="repeated run differs")
        out = lab
    msgs = plan.dump_msgs(dump_level) if dump_level is not None else None
    return out, msgs, plan


def gpu_energy(data: np.ndarray, table: np.ndarray, labels: np.ndarray) -> float:
    torch = torch_cuda()
    L, H, W = data.shape
    dd = torch.from_numpy(np.ascontiguousarray(data)).cuda()
    td = torch.from_numpy(np.ascontiguousarray(table, dtype=data.dtype)).cuda()
    ld = torch.from_numpy(np.ascontiguousarray(labels, dtype=np.int32)).cuda()
    return G.energy_f64(W, H, L, G.dtype_code(data.dtype), dd.data_ptr(), td.data_ptr(), ld.data_ptr(),
                        torch.cuda.current_stream().cuda_stream)
