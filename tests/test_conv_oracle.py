"""Pins for the convolution oracle (oracle/conv_oracle.py), CPU only:
brute-force loops for the forward, central finite differences for every
analytic gradient, and the bf16 rounding helper against exact cases."""
import numpy as np
import pytest

from oracle import conv_oracle as C


def brute_conv(x, w, b):
    N, H, W, Ci = x.shape
    Co = w.shape[0]
    y = np.zeros((N, H, W, Co))
    for n in range(N):
        for h in range(H):
            for ww in range(W):
                for co in range(Co):
                    acc = b[co]
                    for kh in range(3):
                        for kw in range(3):
                            hh, w2 = h + kh - 1, ww + kw - 1
                            if 0 <= hh < H and 0 <= w2 < W:
                                acc += np.dot(w[co, kh, kw], x[n, hh, w2])
                    y[n, h, ww, co] = acc
    return y


def test_conv_forward_brute():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 4, 5, 3))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    np.testing.assert_allclose(C.conv3x3(x, w, b), brute_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv_backward_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 4, 3, 2))
    w = rng.standard_normal((3, 3, 3, 2))
    b = rng.standard_normal(3)
    dz = rng.standard_normal((1, 4, 3, 3))
    dx, dw, db = C.conv3x3_bwd(x, w, dz)
    f = lambda xx, ww, bb: float((C.conv3x3(xx, ww, bb) * dz).sum())
    eps = 1e-6
    for arr, grad, which in ((x, dx, 0), (w, dw, 1), (b, db, 2)):
        for idx in np.ndindex(arr.shape):
            a1 = arr.copy(); a1[idx] += eps
            a2 = arr.copy(); a2[idx] -= eps
            args1 = [x, w, b]; args1[which] = a1
            args2 = [x, w, b]; args2[which] = a2
            num = (f(*args1) - f(*args2)) / (2 * eps)
            assert abs(num - grad[idx]) < 1e-6 * max(1.0, abs(num)), (which, idx)


def test_net_backward_finite_differences():
    rng = np.random.default_rng(2)
    chans = [3, 4, 4, 5]
    ws = [rng.standard_normal((chans[k + 1], 3, 3, chans[k])) * 0.5 for k in range(3)]
    bs = [rng.standard_normal(chans[k + 1]) * 0.1 for k in range(3)]
    x = rng.standard_normal((1, 3, 4, 3))
    t = rng.standard_normal((1, 3, 4, 5))

    def loss(wsx):
        out, _, _ = C.net_forward(x, wsx, bs, emulate_bf16=False)
        return float(((out - t) ** 2).sum())

    out, ins, outs = C.net_forward(x, ws, bs, emulate_bf16=False)
    grads, _ = C.net_backward(2 * (out - t), ws, ins, outs, emulate_bf16=False)
    eps = 1e-6
    for k in range(3):
        for idx in [(0, 0, 0, 0), (1, 2, 1, 0), tuple(s - 1 for s in ws[k].shape)]:
            w1 = [w.copy() for w in ws]; w1[k][idx] += eps
            w2 = [w.copy() for w in ws]; w2[k][idx] -= eps
            num = (loss(w1) - loss(w2)) / (2 * eps)
            assert abs(num - grads[k][0][idx]) < 1e-5 * max(1.0, abs(num)), (k, idx)


def test_bf16_round_cases():
    assert C.bf16_round(1.0) == 1.0
    # 1 + 2^-8 is exactly halfway between bf16 neighbours 1 and 1+2^-7: ties to even -> 1
    assert C.bf16_round(1.0 + 2 ** -8) == 1.0
    assert C.bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert C.bf16_round(-2.5) == -2.5
    bits = C.to_bf16_bits(np.array([1.0, -2.0, 0.5]))
    np.testing.assert_array_equal(C.from_bf16_bits(bits), [1.0, -2.0, 0.5])
