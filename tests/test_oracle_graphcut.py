"""Pins for the oracle's alpha-beta swap (CPU only).

One move is checked against brute force over every relabelling of the pair's
nodes: its full-grid energy (all data terms, every 4-neighbour edge once) is
the minimum, and its alpha set is the smallest optimal one (the nodes that
every minimum cut puts on the sink side), which fixes the labels among ties.
The decoder keeps only moves that lower getLabelingCost and stops at a
labelling no single move improves."""
import itertools

import numpy as np
import pytest

DIRS = [(0, 1), (1, 0), (0, -1), (-1, 0)]


def full_energy(data, table, f):
    """sum over all nodes of D(f) + sum over undirected 4-edges of V(f_u, f_v)."""
    L, H, W = data.shape
    e = int(np.take_along_axis(data, f[None], 0).sum())
    e += int(table[f[:, 1:], f[:, :-1]].sum())   # V symmetric: orientation irrelevant
    e += int(table[f[1:, :], f[:-1, :]].sum())
    return e


def sym_table(rng, L, kind):
    a = np.arange(L)
    if kind == "potts":
        return (7 * (a[:, None] != a[None, :])).astype(np.int32)
    if kind == "tl":
        return np.minimum(3 * np.abs(a[:, None] - a[None, :]), 8).astype(np.int32)
    t = rng.integers(0, 12, size=(L, L))
    t = np.triu(t, 1)
    return (t + t.T).astype(np.int32)  # symmetric, zero diagonal


@pytest.mark.parametrize("kind", ["potts", "tl", "random"])
@pytest.mark.parametrize("seed", range(6))
def test_move_is_optimal_and_minimal(oracle_mod, kind, seed):
    rng = np.random.default_rng(seed * 31 + len(kind))
    L = 4
    H, W = int(rng.integers(2, 5)), int(rng.integers(2, 5))
    data = rng.integers(0, 15, size=(L, H, W)).astype(np.int32)
    table = sym_table(rng, L, kind)
    f = rng.integers(0, L, size=(H, W)).astype(np.int32)
    for alpha, beta in [(0, 1), (1, 3), (0, 2)]:
        got = oracle_mod.ab_move(data, table, f, alpha, beta)
        idx = [tuple(p) for p in np.argwhere((f == alpha) | (f == beta))]
        outside = (f != alpha) & (f != beta)
        np.testing.assert_array_equal(got[outside], f[outside])
        best, optimal = None, []
        for bits in itertools.product((0, 1), repeat=len(idx)):
            g = f.copy()
            for (y, x), b in zip(idx, bits):
                g[y, x] = alpha if b else beta
            e = full_energy(data, table, g)
            if best is None or e < best:
                best, optimal = e, [g]
            elif e == best:
                optimal.append(g)
        assert full_energy(data, table, got) == best
        got_alpha = (got == alpha) & ~outside
        for g in optimal:
            assert not (got_alpha & ~(g == alpha)).any(), "alpha set not contained in an optimal one"


@pytest.mark.parametrize("kind", ["potts", "tl"])
def test_decoder_descends_to_swap_local_optimum(oracle_mod, kind):
    rng = np.random.default_rng(7)
    L, H, W = 5, 9, 11
    data = rng.integers(0, 20, size=(L, H, W)).astype(np.int32)
    table = sym_table(rng, L, kind)
    init = rng.integers(0, L, size=(H, W)).astype(np.int32)
    e0 = oracle_mod.energy(data, table, init)
    out, moves = oracle_mod.alpha_beta(data, table, init, 50)
    e1 = oracle_mod.energy(data, table, out)
    assert moves > 0 and e1 < e0
    # converged: no single move lowers the energy any more
    for a in range(L):
        for b in range(a + 1, L):
            assert oracle_mod.energy(data, table, oracle_mod.ab_move(data, table, out, a, b)) >= e1
    # zero iterations leave the labelling alone
    same, n0 = oracle_mod.alpha_beta(data, table, init, 0)
    np.testing.assert_array_equal(same, init)
    assert n0 == 0


# ---- the reference's own alpha-beta (graph state kept across pairs) --------

def _tiny_case(rng, kind):
    L = int(rng.integers(2, 5))
    H, W = int(rng.integers(1, 6)), int(rng.integers(1, 6))
    data = rng.integers(0, 10, size=(L, H, W)).astype(np.int32)
    if kind == "asym":
        table = rng.integers(0, 8, size=(L, L)).astype(np.int32)
    else:
        table = sym_table(rng, L, kind)
    init = rng.integers(0, L, size=(H, W)).astype(np.int32)
    return data, table, init


@pytest.mark.parametrize("kind", ["potts", "tl", "random", "asym"])
@pytest.mark.parametrize("seed", range(10))
def test_ref_alpha_beta_matches_python_restatement(oracle_mod, kind, seed):
    """oracle_alpha_beta_ref (C) equals an independent dictionary-based Python
    restatement of src/graphcut.cc on tiny grids: labels and kept moves."""
    from graphcut_ref_py import alpha_beta_ref
    rng = np.random.default_rng(1000 + 17 * seed + len(kind))
    data, table, init = _tiny_case(rng, kind)
    iters = int(rng.integers(1, 4))
    got, kept = oracle_mod.alpha_beta_ref(data, table, init, iters)
    want, want_kept = alpha_beta_ref(data.tolist(), table.tolist(), init.tolist(), iters)
    np.testing.assert_array_equal(got, np.array(want, np.int32))
    assert kept == want_kept


@pytest.mark.parametrize("kind", ["potts", "tl", "random"])
@pytest.mark.parametrize("seed", range(8))
def test_ref_alpha_beta_two_labels_is_the_fresh_swap(oracle_mod, kind, seed):
    """With two labels every node is in the one pair, so graphCutConstructGraph
    rebuilds the whole graph state for it and the reference's decode must
    equal the fresh-graph oracle's (whose moves are pinned by brute force
    above): same labels, same kept moves."""
    rng = np.random.default_rng(seed * 5 + len(kind))
    H, W = int(rng.integers(2, 9)), int(rng.integers(2, 9))
    data = rng.integers(0, 20, size=(2, H, W)).astype(np.int32)
    table = sym_table(rng, 2, kind)
    init = rng.integers(0, 2, size=(H, W)).astype(np.int32)
    got, kept = oracle_mod.alpha_beta_ref(data, table, init, 3)
    want, want_kept = oracle_mod.alpha_beta(data, table, init, 3)
    np.testing.assert_array_equal(got, want)
    assert kept == want_kept


def test_ref_alpha_beta_descends_and_differs_from_fresh(oracle_mod):
    """The reference's decode never raises getLabelingCost and repeats itself
    given enough rounds; because the state of nodes outside a pair carries
    over to the next pair, it is not the fresh-graph decode: among these
    small random grids (4 % of such grids, in a search over 20,000) some end
    with other labels."""
    rng = np.random.default_rng(1)
    differ = 0
    for k in range(60):
        L, H, W = int(rng.integers(3, 6)), int(rng.integers(3, 7)), int(rng.integers(3, 7))
        data = rng.integers(0, int(rng.integers(2, 30)), size=(L, H, W)).astype(np.int32)
        kind = ["potts", "tl", "random", "asym"][k % 4]
        table = rng.integers(0, 10, size=(L, L)).astype(np.int32) if kind == "asym" else sym_table(rng, L, kind)
        init = rng.integers(0, L, size=(H, W)).astype(np.int32)
        e0 = oracle_mod.energy(data, table, init)
        got, kept = oracle_mod.alpha_beta_ref(data, table, init, 3)
        e1 = oracle_mod.energy(data, table, got)
        assert e1 <= e0 and (kept > 0) == (e1 < e0)
        fresh, _ = oracle_mod.alpha_beta(data, table, init, 3)
        differ += int(not np.array_equal(fresh, got))
    assert differ >= 1
