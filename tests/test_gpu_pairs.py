"""Fused pairs of half-sweeps (two half-sweeps per launch, the first colour's
messages passed to the second through LDS) against the oracle: labels and
messages bit for bit, for every label capacity the pair kernel is built for,
odd and even half-sweep counts, several levels, both storage forms of the
data costs, seeded schedules, and the headline workload at full size."""
import zlib

import numpy as np
import pytest

from gpu_util import run_plan, gpu_energy, say, torch_cuda
from gridinfcuda_amd import _lib as G

pytestmark = pytest.mark.gpu

SM = dict(potts=1, tl=2)


def _table(o, kind, L, dtype):
    return o.table_potts(L, 15, dtype) if kind == "potts" else o.table_trunc_linear(L, 8, 32, dtype)


def _cases():
    rng = np.random.default_rng(77)
    shapes = [(3, 3), (4, 5), (9, 7), (17, 33), (40, 41), (23, 70), (61, 37), (5, 130), (130, 6)]
    Ls = [2, 5, 16, 17, 31, 32, 33, 64, 65, 96, 100, 128]
    out = []
    for k in range(48):
        H, W = shapes[k % len(shapes)]
        L = Ls[k % len(Ls)]
        levels = int(rng.integers(1, 4))
        iters = int(rng.integers(1, 8))
        kind = ["tl", "potts"][k % 2]
        dt, norm = [("float32", 0), ("int32", 0), ("float32", 1), ("int32", 1)][(k // 2) % 4]
        out.append((H, W, L, levels, iters, kind, dt, norm))
    return out


@pytest.mark.parametrize("H,W,L,levels,iters,kind,dt,norm", _cases())
def test_pairs_bit_exact(oracle_mod, H, W, L, levels, iters, kind, dt, norm):
    say(f"pairs H={H} W={W} L={L} levels={levels} iters={iters} {kind} {dt} norm={norm}")
    rng = np.random.default_rng(zlib.crc32(repr((H, W, L, levels, iters, kind, dt, norm)).encode()))
    dtype = np.dtype(dt)
    data = rng.integers(0, 40, size=(L, H, W)).astype(dtype)
    table = _table(oracle_mod, kind, L, dtype)
    init = rng.integers(0, L, size=(H, W)).astype(np.int32)
    # the data-cost storage forms in turn: narrowest that fits (8-bit here), at most 16-bit, T
    lab, msgs, plan = run_plan(data, table, SM[kind], levels, iters, norm=norm, labels_init=init, dump_level=0,
                               kernel=2, compact=[1, 2, 0][(H + W + L) % 3], pairs=True)
    assert plan.pairs() == (L <= 128)
    ref_lab, ref_msgs, _ = oracle_mod.hbp(data, table, SM[kind], levels, iters, norm=norm, labels_init=init,
                                          return_state=True)
    np.testing.assert_array_equal(lab, ref_lab)
    np.testing.assert_array_equal(msgs, ref_msgs)


@pytest.mark.parametrize("L", [16, 32, 64, 96, 128])
def test_pairs_stereo_levels_equal_unfused(oracle_mod, L):
    """A stereo instance spanning many tiles at every level (643x389, 4
    levels x 9 half-sweeps: an odd count, so the last half-sweep of each level
    runs unfused): labels and the level-0 and level-1 messages of the fused
    run equal the unfused run's and the oracle's labels."""
    say(f"pairs stereo 643x389x{L}")
    g = G.gen_stereo(11 + L, 643, 389, L, 24, L - 1, 40, want=("cost_f32",))
    data = g["cost_f32"]
    table = oracle_mod.table_trunc_linear(L, 8, 32)
    lab_u, msgs_u, plan_u = run_plan(data, table, 2, 4, 9, norm=0, dump_level=0, kernel=2, pairs=False)
    m1_u = plan_u.dump_msgs(1)
    del plan_u
    lab_p, msgs_p, plan_p = run_plan(data, table, 2, 4, 9, norm=0, dump_level=0, kernel=2, pairs=True, runs=2)
    assert plan_p.pairs()
    np.testing.assert_array_equal(lab_p, lab_u)
    assert np.array_equal(msgs_p, msgs_u), f"{int((msgs_p != msgs_u).sum())} level-0 messages differ"
    assert np.array_equal(plan_p.dump_msgs(1), m1_u)
    ref = oracle_mod.hbp(data, table, 2, 4, 9, norm=0, threads=16)
    np.testing.assert_array_equal(lab_p, ref)


def test_pairs_seeded_schedule():
    """A seeded scheduled run (the warm start's entry) is the same with and
    without pairs."""
    torch = torch_cuda()
    say("pairs seeded schedule")
    W, H, L = 150, 101, 64
    data = G.gen_stereo(5, W, H, L, 10, L - 1, 40, want=("cost_f32",))["cost_f32"]
    a = np.arange(L)
    table = np.minimum(10 * np.abs(a[:, None] - a[None, :]), 30).astype(np.float32)
    s = torch.cuda.current_stream().cuda_stream
    dd = torch.from_numpy(data).cuda()
    rng = np.random.default_rng(3)
    pred = torch.from_numpy(rng.standard_normal(((H + 3) // 4, (W + 3) // 4, 4 * L)).astype(np.float32)).cuda()
    out = []
    for pairs in (False, True):
        plan = G.Plan(G.make_desc(W, H, L, G.F32, G.SM_TRUNC_LINEAR, 3, 10), table)
        plan.set_pairs(pairs)
        lab = torch.zeros((H, W), dtype=torch.int32, device="cuda")
        plan.seed(1, pred.data_ptr(), pred.shape[1], pred.shape[0], 1, 0.5, s)
        plan.run_schedule(dd.data_ptr(), lab.data_ptr(), 1, [7, 4], True, s)
        torch.cuda.synchronize()
        out.append((lab.cpu().numpy(), plan.dump_msgs(0), plan.dump_msgs(1)))
    for x, y in zip(out[0], out[1]):
        np.testing.assert_array_equal(x, y)


def test_pairs_config2_full_size(oracle_mod):
    """The headline workload (1920x1080x128, seed 2, TL lambda=8 tau=4, 5
    levels x 10 half-sweeps, FP32 mean normalisation) with fused pairs:
    labels, level-0 messages and FP64 energy equal the oracle's."""
    say("pairs config2 full 1920x1080x128")
    data = G.gen_stereo(2, 1920, 1080, 128, 64, 127, 40, want=("cost_f32",))["cost_f32"]
    table = oracle_mod.table_trunc_linear(128, 8, 32)
    ref_lab, ref_msgs, _ = oracle_mod.hbp(data, table, 2, 5, 10, norm=0, threads=16, return_state=True)
    lab, msgs, plan = run_plan(data, table, 2, 5, 10, norm=0, dump_level=0, pairs=True)
    assert plan.pairs()
    np.testing.assert_array_equal(lab, ref_lab)
    assert np.array_equal(msgs, ref_msgs), f"{int((msgs != ref_msgs).sum())} messages differ"
    assert gpu_energy(data, table, lab) == oracle_mod.energy_f64(data, table, ref_lab)
