"""Pins for the oracle's ICM and max-prior restatements (CPU only).

decodeMaxPrior against numpy's first-occurrence argmin; one synchronous ICM
iteration against a vectorised numpy evaluation of the local cost with an
asymmetric table (a transposed V or a wrong neighbour offset fails it);
asynchronous ICM by the energy descent it guarantees for symmetric tables,
by local optimality of every interior node at its fixed point, and on a
single interior node with an asymmetric table."""
import numpy as np
import pytest

# neighbour offsets in the reference direction order S, E, N, W (dx, dy)
DIRS = [(0, 1), (1, 0), (0, -1), (-1, 0)]


def local_costs(data, table, f):
    """cost[l, y, x] = D(x,y,l) + sum_d V(l, f(x+dx, y+dy)) for interior nodes,
    with V(a, b) = table[b, a] (the Matrix layout a + L*b)."""
    L, H, W = data.shape
    c = data[:, 1:H - 1, 1:W - 1].astype(np.int64)
    for dx, dy in DIRS:
        nb = f[1 + dy:H - 1 + dy, 1 + dx:W - 1 + dx]
        c = c + table[nb, :].transpose(2, 0, 1).astype(np.int64)
    return c


def rand_case(rng, L, H, W, dtype, asym=True):
    data = rng.integers(0, 25, size=(L, H, W)).astype(dtype)
    if asym:
        table = rng.integers(0, 15, size=(L, L)).astype(dtype)
    else:
        a = np.arange(L)
        table = np.minimum(4 * np.abs(a[:, None] - a[None, :]), 11).astype(dtype)
    init = rng.integers(0, L, size=(H, W)).astype(np.int32)
    return data, table, init


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
def test_max_prior_first_argmin(oracle_mod, dtype):
    rng = np.random.default_rng(1)
    data = rng.integers(0, 4, size=(6, 9, 11)).astype(dtype)  # many ties
    np.testing.assert_array_equal(oracle_mod.max_prior(data), np.argmin(data, axis=0))


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
@pytest.mark.parametrize("L,H,W", [(5, 7, 9), (8, 3, 12), (3, 10, 4)])
def test_icm_sync_one_step_matches_numpy(oracle_mod, dtype, L, H, W):
    rng = np.random.default_rng(L * 100 + H)
    data, table, init = rand_case(rng, L, H, W, dtype)
    want = init.copy()
    f = init.copy()
    for _ in range(3):
        nxt = f.copy()
        nxt[1:H - 1, 1:W - 1] = np.argmin(local_costs(data, table, f), axis=0)
        f = nxt
        if _ == 0:
            want = f.copy()
    np.testing.assert_array_equal(oracle_mod.icm(data, table, init, 1, sync=True), want)
    np.testing.assert_array_equal(oracle_mod.icm(data, table, init, 3, sync=True), f)
    np.testing.assert_array_equal(oracle_mod.icm(data, table, init, 0, sync=True), init)


def test_icm_without_smoothness_is_max_prior_inside(oracle_mod):
    rng = np.random.default_rng(2)
    data, _, init = rand_case(rng, 7, 8, 9, np.int32)
    zero = np.zeros((7, 7), np.int32)
    mp = oracle_mod.max_prior(data)
    for sync in (True, False):
        got = oracle_mod.icm(data, zero, init, 2, sync=sync)
        np.testing.assert_array_equal(got[1:-1, 1:-1], mp[1:-1, 1:-1])
        border = np.ones_like(init, bool)
        border[1:-1, 1:-1] = False
        np.testing.assert_array_equal(got[border], init[border])


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
def test_icm_async_descends_to_local_optimum(oracle_mod, dtype):
    rng = np.random.default_rng(3)
    data, table, init = rand_case(rng, 6, 11, 13, dtype, asym=False)
    f = init
    e_prev = oracle_mod.energy_f64(data, table, f)
    for it in range(1, 30):
        g = oracle_mod.icm(data, table, init, it, sync=False)
        e = oracle_mod.energy_f64(data, table, g)
        assert e <= e_prev
        e_prev = e
        if np.array_equal(g, f):
            break
        f = g
    else:
        pytest.fail("async ICM did not reach a fixed point")
    # at the fixed point every interior label is the first argmin of its local cost
    np.testing.assert_array_equal(f[1:-1, 1:-1], np.argmin(local_costs(data, table, f), axis=0))


def test_icm_single_interior_node_asymmetric(oracle_mod):
    rng = np.random.default_rng(4)
    for _ in range(20):
        data, table, init = rand_case(rng, 9, 3, 3, np.int32)
        want = init.copy()
        want[1, 1] = int(np.argmin(local_costs(data, table, init)[:, 0, 0]))
        for sync in (True, False):
            np.testing.assert_array_equal(oracle_mod.icm(data, table, init, 1, sync=sync), want)


def test_icm_async_row_order(oracle_mod):
    """A 1-row interior (H=3): node x sees the new label of x-1 and the old
    label of x+1; computed here node by node with numpy local costs."""
    rng = np.random.default_rng(5)
    data, table, init = rand_case(rng, 5, 3, 9, np.int32)
    f = init.copy()
    for x in range(1, 8):
        c = data[:, 1, x].astype(np.int64)
        for dx, dy in DIRS:
            c = c + table[f[1 + dy, x + dx], :]
        f[1, x] = int(np.argmin(c))
    np.testing.assert_array_equal(oracle_mod.icm(data, table, init, 1, sync=False), f)
