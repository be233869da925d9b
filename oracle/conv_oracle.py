"""ORACLE -- test infrastructure only.  Float64 numpy restatement of the
warm-start network's layers (3x3 convolution, zero padding 1, NHWC, weights
[Cout][3][3][Cin]) with the BF16 rounding points of the GPU path made
explicit, so GPU forward/backward results can be compared against it.  The
analytic gradients here are pinned by central finite differences
(tests/test_conv_oracle.py).  The warm start has no reference counterpart.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a) -> np.ndarray:
    """Round float values to the nearest bfloat16 (ties to even), as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def to_bf16_bits(a) -> np.ndarray:
    f = np.asarray(bf16_round(a), np.float32)
    return (f.view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(b) -> np.ndarray:
    return (np.asarray(b, np.uint32) << 16).view(np.float32).astype(np.float64)


def conv3x3(x, w, b=None):
    """y[n,h,w,co] = b[co] + sum_{kh,kw,ci} w[co,kh,kw,ci] x[n,h+kh-1,w+kw-1,ci]."""
    N, H, W, Ci = x.shape
    Co = w.shape[0]
    xp = np.zeros((N, H + 2, W + 2, Ci))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((N, H, W, Co))
    for kh in range(3):
        for kw in range(3):
            y += xp[:, kh:kh + H, kw:kw + W, :] @ w[:, kh, kw, :].T
    if b is not None:
        y += b
    return y


def conv3x3_bwd(x, w, dz):
    """Gradients of sum(dz * conv3x3(x, w, b)): (dx, dw, db)."""
    N, H, W, Ci = x.shape
    xp = np.zeros((N, H + 2, W + 2, Ci))
    xp[:, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w, dtype=np.float64)
    for kh in range(3):
        for kw in range(3):
            sl = xp[:, kh:kh + H, kw:kw + W, :]
            dw[:, kh, kw, :] = np.einsum("nhwo,nhwi->oi", dz, sl)
            dxp[:, kh:kh + H, kw:kw + W, :] += dz @ w[:, kh, kw, :]
    db = dz.sum(axis=(0, 1, 2))
    return dxp[:, 1:-1, 1:-1, :], dw, db


def net_forward(x, weights, biases, emulate_bf16=True):
    """Layers conv-ReLU ... conv (no ReLU on the last).  With emulate_bf16 the
    layer inputs and weights are rounded to bf16 as the GPU stores them.
    Returns (output, list of layer inputs, list of post-activation outputs)."""
    r = bf16_round if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    h = r(x)
    inputs, outs = [], []
    for k, (w, b) in enumerate(zip(weights, biases)):
        inputs.append(h)
        z = conv3x3(h, r(w), b)
        last = k == len(weights) - 1
        a = z if last else np.maximum(z, 0.0)
        outs.append(a)
        h = a if last else r(a)
    return outs[-1], inputs, outs


def net_backward(dout, weights, inputs, outs, emulate_bf16=True):
    """Backprop through net_forward; dz of each layer rounded to bf16 like the GPU."""
    r = bf16_round if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    grads = [None] * len(weights)
    dy = dout
    for k in range(len(weights) - 1, -1, -1):
        last = k == len(weights) - 1
        dz = dy if last else dy * (r(outs[k]) > 0)
        dz = r(dz)
        dx, dw, db = conv3x3_bwd(inputs[k], r(weights[k]), dz)
        grads[k] = (dw, db)
        dy = dx
    return grads, dy
