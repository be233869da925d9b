"""Independent slow restatements used to pin the oracle (test helper).

These are written from the mathematics of min-sum BP rather than from the
oracle's loop structure: messages live in a dict keyed by directed edge,
the distance transform is the brute-force min-convolution
``min_j(h_j + V(i, j))`` over the full smoothness table, and the red-black
schedule is expressed as "which colour moves at half-sweep i".  Pure Python,
so only tiny grids.
"""
from __future__ import annotations

import numpy as np

# (dx, dy) per direction in the reference order S, E, N, W
# (reference src/include/utils.h:26-37).
DIRS = [(0, 1), (1, 0), (0, -1), (-1, 0)]


def _interior(W, H):
    return [(x, y) for y in range(1, H - 1) for x in range(1, W - 1)]


def brute_bp(data, table, model, iters, norm="mean", levels=1, labels_init=None, dtype=None):
    """Min-sum BP/HBP on a 4-connected grid.

    data: (L, H, W) costs; table: (L, L) with table[b, a] = V(a, b).
    model: "linear" (TL/Potts path: reverse message subtracted, normalised)
           or "general" (plain min-convolution of the belief).
    Arithmetic is done with Python ints for integer dtypes (C truncating
    division for the mean) and numpy float32 scalars otherwise.
    """
    dtype = np.dtype(dtype or data.dtype)
    is_int = dtype.kind == "i"
    L, H, W = data.shape

    def cast(v):
        return int(v) if is_int else np.float32(v)

    def V(a, b):
        return cast(table[b, a])

    # Pyramid: level dims and summed costs (parents own 2x2 children).
    dims = [(W, H)]
    costs = [{(x, y): [cast(data[l, y, x]) for l in range(L)] for y in range(H) for x in range(W)}]
    for _ in range(1, levels):
        w, h = dims[-1]
        cw, ch = -(-w // 2), -(-h // 2)
        c = {}
        for Y in range(ch):
            for X in range(cw):
                acc = [cast(0)] * L
                for (x, y) in [(2 * X, 2 * Y), (2 * X + 1, 2 * Y), (2 * X, 2 * Y + 1), (2 * X + 1, 2 * Y + 1)]:
                    if x < w and y < h:
                        acc = [acc[l] + costs[-1][(x, y)][l] for l in range(L)]
                c[(X, Y)] = acc
        dims.append((cw, ch))
        costs.append(c)

    out_msgs = None   # dict (node, dir) -> list[L]; node's outbox towards dir
    belief = None
    for t in range(levels - 1, -1, -1):
        w, h = dims[t]
        msgs = {}
        for y in range(h):
            for x in range(w):
                for d in range(4):
                    if out_msgs is None:
                        msgs[((x, y), d)] = [cast(0)] * L
                    else:
                        msgs[((x, y), d)] = list(out_msgs[((x // 2, y // 2), d)])
        belief = {p: [cast(0)] * L for y in range(h) for x in range(w) for p in [(x, y)]}
        for i in range(iters):
            movers = [(x, y) for (x, y) in _interior(w, h) if (x + y + i) % 2 == 1]
            for (x, y) in movers:
                incoming = []
                for d, (dx, dy) in enumerate(DIRS):
                    q = (x + dx, y + dy)
                    back = (d + 2) % 4
                    incoming.append(msgs[(q, back)])
                b = []
                for l in range(L):
                    v = costs[t][(x, y)][l]
                    for d in range(4):
                        v = v + incoming[d][l]
                    b.append(v)
                belief[(x, y)] = b
                new = []
                for d in range(4):
                    if model == "linear":
                        hvec = [b[l] - incoming[d][l] for l in range(L)]
                        m = [min(hvec[j] + V(k, j) for j in range(L)) for k in range(L)]
                        if norm == "mean":
                            s = cast(0)
                            for v in m:
                                s = s + v
                            if is_int:
                                # C division truncates toward zero.
                                nu = abs(s) // L if s >= 0 else -(abs(s) // L)
                            else:
                                nu = np.float32(s / np.float32(L))
                        else:
                            nu = min(m)
                        m = [v - nu for v in m]
                    else:
                        m = [min(V(k, j) + b[j] for j in range(L)) for k in range(L)]
                    new.append(m)
                for d in range(4):
                    msgs[((x, y), d)] = new[d]
        out_msgs = msgs
    labels = (np.zeros((H, W), np.int32) if labels_init is None else np.array(labels_init, np.int32))
    for (x, y) in _interior(W, H):
        b = belief[(x, y)]
        labels[y, x] = int(min(range(L), key=lambda l: (b[l], l)))
    return labels, out_msgs, belief


def viterbi_chain(costs_row, table):
    """Exact MAP of a chain: sum_x D(x, f_x) + sum_x V(f_x, f_{x+1}).

    costs_row: (L, n) costs of the chain nodes; table[b, a] = V(a, b).
    Returns (labels, energy) computed in float64 by dynamic programming."""
    L, n = costs_row.shape
    Vm = np.asarray(table, np.float64).T  # Vm[a, b] = V(a, b)
    acc = np.asarray(costs_row[:, 0], np.float64)
    back = []
    for x in range(1, n):
        tot = acc[:, None] + Vm  # [prev, cur]
        arg = np.argmin(tot, axis=0)
        back.append(arg)
        acc = tot[arg, np.arange(L)] + costs_row[:, x]
    f = [int(np.argmin(acc))]
    for arg in reversed(back):
        f.append(int(arg[f[-1]]))
    f.reverse()
    return np.array(f), float(np.min(acc))
