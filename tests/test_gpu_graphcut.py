"""GPU alpha-beta swap against the oracle: every kept move is the canonical
optimal swap (the sink-reachable set of any maximum flow takes alpha), so the
whole decode -- labels and the number of kept moves -- is reproduced bit for
bit."""
import numpy as np
import pytest

from gpu_util import torch_cuda, say
from gridinfcuda_amd import _lib as G

pytestmark = pytest.mark.gpu


def sym_table(rng, L, kind):
    a = np.arange(L)
    if kind == "potts":
        return (7 * (a[:, None] != a[None, :])).astype(np.int32)
    if kind == "tl":
        return np.minimum(3 * np.abs(a[:, None] - a[None, :]), 8).astype(np.int32)
    t = np.triu(rng.integers(0, 12, size=(L, L)), 1)
    return (t + t.T).astype(np.int32)


@pytest.mark.parametrize("kind", ["potts", "tl", "random"])
@pytest.mark.parametrize("L,H,W,iters", [(3, 3, 3, 2), (4, 9, 13, 3), (6, 24, 32, 4), (5, 2, 17, 2), (8, 31, 7, 2),
                                         (1, 5, 5, 3), (4, 12, 12, 0), (3, 1, 1, 2), (4, 1, 9, 2), (3, 2, 2, 2),
                                         (4, 7, 2, 2)])
def test_alpha_beta_matches_oracle(oracle_mod, kind, L, H, W, iters):
    """Includes grids without interior nodes (1x1, 1xN, 2x2, Nx2): their
    labelling cost is 0 for every labelling, so no move is ever kept and the
    initial labels come back unchanged."""
    torch = torch_cuda()
    say(f"alpha-beta {kind} L={L} {W}x{H} iters={iters}")
    rng = np.random.default_rng(L * 1000 + H * 10 + W)
    data = rng.integers(0, 25, size=(L, H, W)).astype(np.int32)
    table = sym_table(rng, L, kind)
    init = rng.integers(0, L, size=(H, W)).astype(np.int32)
    want, want_moves = oracle_mod.alpha_beta(data, table, init, iters)
    dd, td, idev = (torch.from_numpy(a).cuda() for a in (data, table, init))
    out = torch.full((H, W), -1, dtype=torch.int32, device="cuda")
    moves = G.alpha_beta(W, H, L, dd.data_ptr(), td.data_ptr(), idev.data_ptr(), iters, out.data_ptr(),
                         torch.cuda.current_stream().cuda_stream)
    np.testing.assert_array_equal(out.cpu().numpy(), want)
    assert moves == want_moves
    np.testing.assert_array_equal(G.alpha_beta_host(data, table, init, iters), want)
    if W < 3 or H < 3:
        np.testing.assert_array_equal(want, init)
        assert moves == 0


def test_alpha_beta_refines_bp(oracle_mod):
    """BP labels refined by alpha-beta swap: the energy never rises and the
    result equals the oracle's."""
    say("bp -> alpha-beta")
    W, H, L = 40, 30, 6
    g = G.gen_stereo(23, W, H, L, 4, L - 1, 20, want=("cost_i32",))
    data = g["cost_i32"]
    a = np.arange(L)
    table = np.minimum(4 * np.abs(a[:, None] - a[None, :]), 12).astype(np.int32)
    lab = G.hbp_host(data, table, G.SM_TRUNC_LINEAR, 2, 6)
    want, _ = oracle_mod.alpha_beta(data, table, lab, 3)
    got = G.alpha_beta_host(data, table, lab, 3)
    np.testing.assert_array_equal(got, want)
    assert oracle_mod.energy(data, table, got) <= oracle_mod.energy(data, table, lab)


def test_alpha_beta_rejects_float():
    data = np.zeros((3, 4, 4), np.float32)
    with pytest.raises(G.GinfError) as e:
        G.alpha_beta_host(data, np.zeros((3, 3), np.float32), np.zeros((4, 4), np.int32), 1)
    assert e.value.code == 2


def test_alpha_beta_rejects_labels_out_of_range():
    data = np.zeros((3, 5, 5), np.int32)
    table = np.zeros((3, 3), np.int32)
    init = np.zeros((5, 5), np.int32)
    init[2, 2] = 7
    with pytest.raises(G.GinfError) as e:
        G.alpha_beta_host(data, table, init, 1)
    assert e.value.code == 1


def test_alpha_beta_against_reference_order(oracle_mod):
    """The GPU decode (each pair from a fresh graph) against a restatement of
    the reference's own decode (graph state of nodes outside a pair carried
    over, sequential push-relabel, src/graphcut.cc:197-279):
    * the 384x288x16 stereo workload of the alpha-beta benchmark from the
      decodeMaxPrior labelling: identical labels and kept moves;
    * the Tsukuba pair's costs (examples/stereo) from the example's BP
      labels: the GPU energy is not above the reference's; both energies and
      the label agreement are printed."""
    import os
    from gridinfcuda_amd import stereo as S
    say("alpha-beta vs reference order: 384x288x16 stereo")
    W, H, L = 384, 288, 16
    cost = G.gen_stereo(1, W, H, L, 8, 15, 20, want=("cost_i32",))["cost_i32"]
    a = np.arange(L)
    table = np.minimum(10 * np.abs(a[:, None] - a[None, :]), 20).astype(np.int32)
    init = oracle_mod.max_prior(cost)
    ref, ref_moves = oracle_mod.alpha_beta_ref(cost, table, init, 1)
    got = G.alpha_beta_host(cost, table, init, 1)
    e_ref, e_got = oracle_mod.energy(cost, table, ref), oracle_mod.energy(cost, table, got)
    say(f"  energy GPU {e_got}, reference order {e_ref} ({ref_moves} moves kept)")
    np.testing.assert_array_equal(got, ref)
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    left = S.load_gray(os.path.join(golden, "tsukuba_left.png"))
    right = S.load_gray(os.path.join(golden, "tsukuba_right.png"))
    costs, table = S.example_costs(left, right), S.smoothness_table()
    say("alpha-beta vs reference order: Tsukuba from the example's BP labels")
    bp = G.hbp_host(costs, table, G.SM_TRUNC_LINEAR, 1, 30)
    ref, ref_moves = oracle_mod.alpha_beta_ref(costs, table, bp, 1)
    got = G.alpha_beta_host(costs, table, bp, 1)
    e_bp = oracle_mod.energy(costs, table, bp)
    e_ref, e_got = oracle_mod.energy(costs, table, ref), oracle_mod.energy(costs, table, got)
    say(f"  energy BP {e_bp}, GPU alpha-beta {e_got}, reference order {e_ref} ({ref_moves} moves kept); "
        f"labels equal on {(got == ref).mean():.6f} of the nodes")
    assert e_got <= e_ref < e_bp
