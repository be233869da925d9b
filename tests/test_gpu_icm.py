"""GPU parity of ICM (synchronous and asynchronous) and the max-prior
labelling against the oracle, bit for bit, through the C ABI."""
import numpy as np
import pytest

from gpu_util import torch_cuda, say
from gridinfcuda_amd import _lib as G

pytestmark = pytest.mark.gpu


def case(seed, L, H, W, dtype, kind):
    rng = np.random.default_rng(seed)
    data = rng.integers(0, 30, size=(L, H, W)).astype(dtype)
    a = np.arange(L)
    if kind == "explicit":
        table = rng.integers(0, 20, size=(L, L)).astype(dtype)
    elif kind == "potts":
        table = (9 * (a[:, None] != a[None, :])).astype(dtype)
    else:
        table = np.minimum(5 * np.abs(a[:, None] - a[None, :]), 15).astype(dtype)
    init = rng.integers(0, L, size=(H, W)).astype(np.int32)
    return data, table, init


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
@pytest.mark.parametrize("mode", [G.ICM_SYNC, G.ICM_ASYNC])
@pytest.mark.parametrize("L,H,W,iters,kind", [(16, 37, 53, 5, "tl"), (7, 3, 40, 3, "explicit"), (5, 30, 3, 4, "potts"),
                                              (32, 129, 300, 2, "explicit"), (3, 2, 9, 3, "tl"), (9, 17, 11, 0, "tl"),
                                              (64, 90, 70, 6, "potts"), (4, 400, 20, 7, "explicit"),
                                              # table in LDS: full 64 KiB, padded rows; then in global memory
                                              (128, 40, 50, 3, "explicit"), (100, 33, 45, 4, "tl"),
                                              (150, 20, 30, 2, "potts"),
                                              # more interior nodes than one pass of the synchronous grid
                                              (8, 700, 800, 3, "tl"),
                                              # W % 4 == 0: four nodes per thread (one quad holding both
                                              # border columns; a full 128-label table in LDS)
                                              (6, 9, 4, 3, "tl"), (128, 24, 64, 3, "explicit"),
                                              (33, 50, 28, 4, "potts")])
def test_icm_matches_oracle(oracle_mod, dtype, mode, L, H, W, iters, kind):
    torch = torch_cuda()
    say(f"icm mode={mode} {np.dtype(dtype).name} L={L} H={H} W={W} iters={iters} {kind}")
    data, table, init = case(L * 7 + H, L, H, W, dtype, kind)
    want = oracle_mod.icm(data, table, init, iters, sync=(mode == G.ICM_SYNC))
    dd, td, idev = (torch.from_numpy(a).cuda() for a in (data, table, init))
    out = torch.full((H, W), -1, dtype=torch.int32, device="cuda")
    G.icm(W, H, L, G.dtype_code(data.dtype), dd.data_ptr(), td.data_ptr(), idev.data_ptr(), iters, mode,
          out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), want)
    # in place (result == initial) and through the host entry
    G.icm(W, H, L, G.dtype_code(data.dtype), dd.data_ptr(), td.data_ptr(), idev.data_ptr(), iters, mode,
          idev.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(idev.cpu().numpy(), want)
    np.testing.assert_array_equal(G.icm_host(data, table, init, iters, mode), want)


@pytest.mark.parametrize("form", ["fits", "past255", "fraction", "negzero"])
def test_icm_sync_cost_storage(oracle_mod, form):
    """Synchronous ICM with four nodes per thread reads an 8-bit copy of the
    costs when every cost is an integer in [0, 255], else the costs
    themselves (decided on the device); either way the labels equal the
    oracle's."""
    torch = torch_cuda()
    L, H, W, iters = 16, 29, 48, 4
    dtype = np.int32 if form == "past255" else np.float32
    say(f"icm sync cost storage {form}")
    data, table, init = case(11, L, H, W, dtype, "tl")
    if form == "past255":
        data[3, 7, 9] = 300
    elif form == "fraction":
        data[5, 10, 20] += 0.5
    elif form == "negzero":
        data[0, 4, 4] = -0.0
    want = oracle_mod.icm(data, table, init, iters, sync=True)
    dd, td, idev = (torch.from_numpy(a).cuda() for a in (data, table, init))
    out = torch.full((H, W), -1, dtype=torch.int32, device="cuda")
    G.icm(W, H, L, G.dtype_code(data.dtype), dd.data_ptr(), td.data_ptr(), idev.data_ptr(), iters, G.ICM_SYNC,
          out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), want)


@pytest.mark.parametrize("L,H,W", [(12, 21, 36), (32, 25, 40), (128, 20, 32)])
@pytest.mark.parametrize("form,dtype", [("fits", np.float32), ("fits", np.int32), ("past255", np.int32),
                                        ("fraction", np.float32), ("negative", np.int32), ("negative", np.float32),
                                        ("negzero", np.float32)])
def test_icm_sync_table_storage(oracle_mod, form, dtype, L, H, W):
    """Synchronous ICM with four nodes per thread on tables whose entries are
    small integers, past 255, fractional, negative or -0, with the costs read
    from their 8-bit copy and in T: the labels equal the oracle's."""
    torch = torch_cuda()
    iters = 3
    say(f"icm sync table storage {form} {np.dtype(dtype).name} L={L} H={H} W={W}")
    data, table, init = case(17 + L, L, H, W, dtype, "tl")
    if form == "past255":
        table[2, 5] = 300
    elif form == "fraction":
        table[1, L - 1] += 0.5
    elif form == "negative":
        table[L - 1, 0] = -1
    elif form == "negzero":
        table[3, 3] = -0.0
    for cost_form in ("fits", "past255"):
        d = data.copy()
        if cost_form == "past255":
            d[1, 5, 9] = 280
        want = oracle_mod.icm(d, table, init, iters, sync=True)
        dd, td, idev = (torch.from_numpy(a).cuda() for a in (d, table, init))
        out = torch.full((H, W), -1, dtype=torch.int32, device="cuda")
        G.icm(W, H, L, G.dtype_code(d.dtype), dd.data_ptr(), td.data_ptr(), idev.data_ptr(), iters, G.ICM_SYNC,
              out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), want)


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
def test_icm_sync_unaligned_buffers(oracle_mod, dtype):
    """W % 4 == 0 but a result buffer 4 bytes past a 16-byte boundary: the
    one-node-per-thread kernel runs, with the same labels."""
    torch = torch_cuda()
    L, H, W, iters = 24, 31, 44, 3
    say(f"icm sync unaligned {np.dtype(dtype).name} L={L} H={H} W={W}")
    data, table, init = case(5, L, H, W, dtype, "tl")
    want = oracle_mod.icm(data, table, init, iters, sync=True)
    dd, td, idev = (torch.from_numpy(a).cuda() for a in (data, table, init))
    base = torch.full((H * W + 4,), -1, dtype=torch.int32, device="cuda")
    out = base[1:1 + H * W]
    G.icm(W, H, L, G.dtype_code(data.dtype), dd.data_ptr(), td.data_ptr(), idev.data_ptr(), iters, G.ICM_SYNC,
          out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy().reshape(H, W), want)


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
@pytest.mark.parametrize("shape", [(12, 33, 47), (12, 32, 47), (1, 8, 6), (19, 4, 1)])
def test_max_prior_matches_oracle(oracle_mod, dtype, shape):
    """W*H % 4 == 0 runs four nodes per thread (second to fourth shapes)."""
    rng = np.random.default_rng(9)
    data = rng.integers(0, 5, size=shape).astype(dtype)
    np.testing.assert_array_equal(G.max_prior_host(data), oracle_mod.max_prior(data))


def test_icm_refines_bp_labels(oracle_mod):
    """BP labels as ICM's initial labelling (the reference's intended chain):
    asynchronous ICM never raises the energy and equals the oracle."""
    say("bp -> icm")
    W, H, L = 96, 64, 32
    data = G.gen_stereo(17, W, H, L, 8, L - 1, 30)["cost_f32"]
    a = np.arange(L)
    table = np.minimum(6 * np.abs(a[:, None] - a[None, :]), 18).astype(np.float32)
    lab = G.hbp_host(data, table, G.SM_TRUNC_LINEAR, 3, 6)
    refined = G.icm_host(data, table, lab, 5, G.ICM_ASYNC)
    np.testing.assert_array_equal(refined, oracle_mod.icm(data, table, lab, 5, sync=False))
    assert oracle_mod.energy_f64(data, table, refined) <= oracle_mod.energy_f64(data, table, lab)


def test_icm_invalid_arguments():
    data = np.zeros((3, 5, 5), np.int32)
    table = np.zeros((3, 3), np.int32)
    init = np.zeros((5, 5), np.int32)
    with pytest.raises(G.GinfError):
        G.icm_host(data, table, init, -1, G.ICM_SYNC)
    with pytest.raises(G.GinfError):
        G.icm_host(data, table, init, 1, 7)
