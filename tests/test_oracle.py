"""Pins for the CPU oracle (oracle/bp_oracle.c), CPU only.

The reference ships no tests, fixtures or golden vectors and cannot be built
here, so every oracle function is pinned by brute force, closed forms,
known answers and an independent restatement (tests/brute_bp.py), chosen so
that a wrong colour parity, direction order, reverse-message index,
normalisation, upsample frame or labelling rule fails at least one test.
"""
import itertools
import zlib

import numpy as np
import pytest

from brute_bp import brute_bp, viterbi_chain, DIRS

SM = dict(explicit=0, potts=1, tl=2, tq=3)


def rand_costs(rng, L, H, W, dtype, hi=20):
    if np.dtype(dtype).kind == "i":
        return rng.integers(0, hi + 1, size=(L, H, W)).astype(np.int32)
    return rng.integers(0, hi + 1, size=(L, H, W)).astype(np.float32)


def make_table(o, kind, L, dtype, rng=None):
    if kind == "potts":
        return o.table_potts(L, 7, dtype)
    if kind == "tl":
        return o.table_trunc_linear(L, 3, 8, dtype)
    if kind == "tq":
        return o.table_trunc_quad(L, 2, 11, dtype)
    # explicit: random asymmetric table
    return rng.integers(0, 12, size=(L, L)).astype(dtype)


# ---------------------------------------------------------------- pyramid

@pytest.mark.parametrize("dtype", [np.float32, np.int32])
@pytest.mark.parametrize("H,W", [(4, 6), (5, 7), (1, 1), (3, 8), (9, 2)])
def test_pyramid_brute(oracle_mod, dtype, H, W):
    rng = np.random.default_rng(H * 31 + W)
    L = 5
    fine = rand_costs(rng, L, H, W, dtype, hi=1000)
    out = oracle_mod.pyramid(fine)
    ch, cw = -(-H // 2), -(-W // 2)
    assert out.shape == (L, ch, cw)
    ref = np.zeros((L, ch, cw), np.float64)
    for y in range(H):
        for x in range(W):
            ref[:, y // 2, x // 2] += fine[:, y, x]
    np.testing.assert_array_equal(out.astype(np.float64), ref)


def test_pyramid_child_order_float(oracle_mod):
    # Non-associative float sums expose the child order (2X,2Y),(2X+1,2Y),(2X,2Y+1),(2X+1,2Y+1).
    fine = np.array([[[1e8, 1.0], [-1e8, 1.0]]], np.float32)  # (L=1, H=2, W=2)
    out = oracle_mod.pyramid(fine)
    expect = np.float32(np.float32(np.float32(np.float32(0) + np.float32(1e8)) + np.float32(1.0))
                        + np.float32(-1e8)) + np.float32(1.0)
    assert out[0, 0, 0] == expect == np.float32(1.0)


# ---------------------------------------------------------------- upsample

@pytest.mark.parametrize("H,W", [(6, 8), (5, 7), (2, 3)])
def test_upsample_closed_form(oracle_mod, H, W):
    rng = np.random.default_rng(7)
    L = 3
    ch, cw = -(-H // 2), -(-W // 2)
    coarse = rng.standard_normal((L, 4, ch, cw)).astype(np.float32)
    fine = oracle_mod.upsample(coarse, H, W)
    ys, xs = np.arange(H) // 2, np.arange(W) // 2
    np.testing.assert_array_equal(fine, coarse[:, :, ys][:, :, :, xs])


# ---------------------------------------------------------------- distance transform

@pytest.mark.parametrize("kind", ["tl", "potts"])
@pytest.mark.parametrize("dtype,norm", [(np.int32, 0), (np.int32, 1), (np.float32, 1)])
@pytest.mark.parametrize("L", [2, 5, 16])
def test_dt_brute_center_node(oracle_mod, kind, dtype, norm, L):
    """3x3 grid: the centre (x+y even) moves at half-sweep 1. Its border
    neighbours' outboxes are random; every outgoing message must equal the
    brute-force min-convolution of (belief - reverse message), normalised."""
    rng = np.random.default_rng(L * 10 + norm)
    data = rand_costs(rng, L, 3, 3, dtype, hi=30)
    msgs = rand_costs(rng, L * 4, 3, 3, dtype, hi=25).reshape(L, 4, 3, 3)
    msgs = np.ascontiguousarray(msgs)
    table = make_table(oracle_mod, kind, L, dtype)
    belief = np.zeros_like(data)
    before = msgs.copy()
    oracle_mod.half_sweep(data, msgs, belief, table, SM[kind], norm, 1)
    inc = [before[:, (d + 2) % 4, 1 + dy, 1 + dx].astype(np.int64) for d, (dx, dy) in enumerate(DIRS)]
    b = data[:, 1, 1].astype(np.int64) + sum(inc)
    np.testing.assert_array_equal(belief[:, 1, 1], b)
    V = table.T.astype(np.int64)  # V[a, b]
    for d in range(4):
        h = b - inc[d]
        m = np.array([min(h[j] + V[i, j] for j in range(L)) for i in range(L)])
        if norm == 0:
            s = int(m.sum())
            nu = abs(s) // L if s >= 0 else -(abs(s) // L)
        else:
            nu = m.min()
        np.testing.assert_array_equal(msgs[:, d, 1, 1].astype(np.int64), m - nu, err_msg=f"d={d}")
    # nothing else moved
    mask = np.ones((3, 3), bool)
    mask[1, 1] = False
    np.testing.assert_array_equal(msgs[:, :, mask], before[:, :, mask])


@pytest.mark.parametrize("kind", ["tq", "explicit"])
def test_general_path_brute(oracle_mod, kind):
    L = 6
    rng = np.random.default_rng(3)
    data = rand_costs(rng, L, 3, 3, np.int32)
    msgs = np.ascontiguousarray(rand_costs(rng, L * 4, 3, 3, np.int32).reshape(L, 4, 3, 3))
    table = make_table(oracle_mod, kind, L, np.int32, rng)
    belief = np.zeros_like(data)
    before = msgs.copy()
    oracle_mod.half_sweep(data, msgs, belief, table, SM[kind], 0, 1)
    b = belief[:, 1, 1].astype(np.int64)
    V = table.T.astype(np.int64)
    m = np.array([min(V[i, j] + b[j] for j in range(L)) for i in range(L)])
    for d in range(4):
        np.testing.assert_array_equal(msgs[:, d, 1, 1], m)
    assert (before != msgs).any()


# ---------------------------------------------------------------- known answers

@pytest.mark.parametrize("dtype", [np.int32, np.float32])
def test_3x3_equals_max_prior(oracle_mod, dtype):
    """Single interior node, zero border messages: BP = argmin D (decodeMaxPrior,
    reference src/include/estimators.h:10-25, strict < so lowest index wins)."""
    rng = np.random.default_rng(11)
    for trial in range(20):
        L = int(rng.integers(2, 12))
        data = rand_costs(rng, L, 3, 3, dtype, hi=6)
        for kind in ["tl", "potts", "tq"]:
            table = make_table(oracle_mod, kind, L, dtype)
            lab = oracle_mod.bp(data, table, SM[kind], 2 + trial % 3)
            assert lab[1, 1] == int(np.argmin(data[:, 1, 1]))


def test_iters_zero_and_one_colour_parity(oracle_mod):
    rng = np.random.default_rng(5)
    L, H, W = 4, 7, 9
    data = rand_costs(rng, L, H, W, np.int32)
    data[0] += 50  # label 0 is never the data minimum
    table = oracle_mod.table_trunc_linear(L, 1, 2, np.int32)
    init = rng.integers(0, L, size=(H, W)).astype(np.int32)
    lab0 = oracle_mod.bp(data, table, 2, 0, labels_init=init)
    assert (lab0[1:-1, 1:-1] == 0).all()
    lab1 = oracle_mod.bp(data, table, 2, 1, labels_init=init)
    yy, xx = np.mgrid[0:H, 0:W]
    inner = (yy > 0) & (yy < H - 1) & (xx > 0) & (xx < W - 1)
    even = inner & ((xx + yy) % 2 == 0)
    odd = inner & ((xx + yy) % 2 == 1)
    assert (lab1[even] == 0).all()          # never updated: zero belief
    assert (lab1[odd] != 0).all()           # updated at half-sweep 0
    for lab in (lab0, lab1):                # border labels untouched
        np.testing.assert_array_equal(lab[~inner], init[~inner])


@pytest.mark.parametrize("H,W", [(2, 5), (5, 2), (1, 1), (2, 2)])
def test_no_interior_leaves_labels(oracle_mod, H, W):
    rng = np.random.default_rng(1)
    L = 3
    data = rand_costs(rng, L, H, W, np.float32)
    init = rng.integers(0, L, size=(H, W)).astype(np.int32)
    lab = oracle_mod.hbp(data, oracle_mod.table_potts(L, 3), 1, 2, 4, labels_init=init)
    np.testing.assert_array_equal(lab, init)


def test_invalid_arguments(oracle_mod):
    data = np.zeros((2, 3, 3), np.float32)
    with pytest.raises(ValueError):
        oracle_mod.hbp(data, oracle_mod.table_potts(2, 1), 1, 0, 3)


# ---------------------------------------------------------------- independent restatement

CASES = []
for dtype, norm in [(np.int32, "mean"), (np.int32, "min"), (np.float32, "min")]:
    for kind in ["tl", "potts", "tq", "explicit"]:
        for (H, W, L, levels, iters) in [(4, 5, 3, 1, 3), (5, 6, 4, 2, 4), (6, 7, 2, 3, 3), (7, 4, 5, 2, 5),
                                         (5, 5, 3, 1, 0), (6, 6, 3, 2, 1)]:
            CASES.append((np.dtype(dtype).name, norm, kind, H, W, L, levels, iters))


@pytest.mark.parametrize("dtype,norm,kind,H,W,L,levels,iters", CASES)
def test_oracle_vs_independent(oracle_mod, dtype, norm, kind, H, W, L, levels, iters):
    rng = np.random.default_rng(zlib.crc32(repr((dtype, norm, kind, H, W, L, levels, iters)).encode()))
    data = rand_costs(rng, L, H, W, dtype, hi=15)
    table = make_table(oracle_mod, kind, L, np.dtype(dtype), rng)
    model = "linear" if kind in ("tl", "potts") else "general"
    normc = 0 if norm == "mean" else 1
    lab, msgs, bel = oracle_mod.hbp(data, table, SM[kind], levels, iters, norm=normc, return_state=True)
    blab, bmsgs, bbel = brute_bp(data, table, model, iters, norm=norm, levels=levels)
    np.testing.assert_array_equal(lab, blab)
    for (x, y), v in bbel.items():
        np.testing.assert_array_equal(bel[:, y, x], np.array(v, dtype=dtype))
    for ((x, y), d), v in bmsgs.items():
        np.testing.assert_array_equal(msgs[:, d, y, x], np.array(v, dtype=dtype))


def test_float_mean_matches_independent(oracle_mod):
    """FP32 mean mode: the independent restatement uses the same sequential
    float32 operations, so results agree to rounding of the sums."""
    rng = np.random.default_rng(99)
    L, H, W = 6, 6, 7
    data = rng.uniform(0, 10, size=(L, H, W)).astype(np.float32)
    table = oracle_mod.table_trunc_linear(L, 2.0, 5.0)
    lab, msgs, bel = oracle_mod.hbp(data, table, 2, 2, 5, norm=0, return_state=True)
    blab, bmsgs, bbel = brute_bp(data, table, "linear", 5, norm="mean", levels=2)
    np.testing.assert_array_equal(lab, blab)
    for ((x, y), d), v in bmsgs.items():
        np.testing.assert_allclose(msgs[:, d, y, x], np.array(v, np.float32), rtol=1e-5, atol=1e-4)


# ---------------------------------------------------------------- exact MAP on chains

@pytest.mark.parametrize("kind", ["tl", "potts"])
@pytest.mark.parametrize("norm", [0, 1])
@pytest.mark.parametrize("seed", range(4))
def test_chain_converges_to_viterbi(oracle_mod, kind, norm, seed):
    """H=3: the interior is a chain (a tree) whose border messages are 0, so
    min-sum BP converges to exact min-marginals; labels must equal the MAP."""
    rng = np.random.default_rng(seed)
    L, W = 7, 14
    data = rng.uniform(0, 10, size=(L, 3, W)).astype(np.float32)
    table = (oracle_mod.table_trunc_linear(L, 1.5, 4.0) if kind == "tl"
             else oracle_mod.table_potts(L, 3.0))
    lab = oracle_mod.bp(data, table, SM[kind], 60, norm=norm)
    f, _ = viterbi_chain(data[:, 1, 1:W - 1], table)
    np.testing.assert_array_equal(lab[1, 1:W - 1], f)


def test_hbp_h3_equals_bp(oracle_mod):
    rng = np.random.default_rng(4)
    L, W = 5, 20
    data = rand_costs(rng, L, 3, W, np.int32)
    table = oracle_mod.table_trunc_linear(L, 2, 6, np.int32)
    a = oracle_mod.hbp(data, table, 2, 4, 7)
    b = oracle_mod.bp(data, table, 2, 7)
    np.testing.assert_array_equal(a, b)


def test_threads_bit_identical(oracle_mod):
    rng = np.random.default_rng(8)
    L, H, W = 8, 33, 41
    data = rng.uniform(0, 20, size=(L, H, W)).astype(np.float32)
    table = oracle_mod.table_trunc_linear(L, 2.0, 6.0)
    a = oracle_mod.hbp(data, table, 2, 3, 6, return_state=True, threads=1)
    b = oracle_mod.hbp(data, table, 2, 3, 6, return_state=True, threads=4)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


# ---------------------------------------------------------------- energy

@pytest.mark.parametrize("dtype", [np.int32, np.float32])
def test_energy_brute(oracle_mod, dtype):
    rng = np.random.default_rng(2)
    L, H, W = 6, 9, 11
    data = rand_costs(rng, L, H, W, dtype, hi=100)
    table = make_table(oracle_mod, "explicit", L, np.dtype(dtype), rng)
    labels = rng.integers(0, L, size=(H, W)).astype(np.int32)
    e = 0
    for y, x in itertools.product(range(1, H - 1), range(1, W - 1)):
        f = labels[y, x]
        e += int(data[f, y, x]) + int(table[labels[y - 1, x], f]) + int(table[labels[y, x + 1], f])
    assert oracle_mod.energy_f64(data, table, labels) == float(e)
    assert float(oracle_mod.energy(data, table, labels)) == float(e)
