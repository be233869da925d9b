"""Batched BP: every instance of a mixed batch (sizes, label counts, Potts /
truncated linear) equals the oracle run on that instance alone; results do
not depend on the tile shape; graph replays are identical."""
import numpy as np
import pytest

from gpu_util import torch_cuda, say
from gridinfcuda_amd import _lib as G
from gridinfcuda_amd import batch as B

pytestmark = pytest.mark.gpu


def small_instances(n, seed):
    """Instances 0-7 cycle through the label classes 16/32/64/128; 8-11 are
    L = 65 and 96 (the capacity-96 kernel, full and partial) with Potts and
    truncated linear each; the rest draw L at random."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        if i < 8:
            L = B.LABEL_CLASSES[i % 4]
        elif i < 12:
            L = (65, 96, 96, 65)[i - 8]
        else:
            L = int(rng.choice([5, 16, 17, 32, 50, 64, 65, 96, 100, 128]))
        W, H = int(rng.integers(3, 70)), int(rng.integers(3, 50))
        potts = i % 2 == 0
        out.append(dict(index=i, seed=900 + i, W=W, H=H, L=L, model=G.SM_POTTS if potts else G.SM_TRUNC_LINEAR,
                        table=B.table_potts(L, 15) if potts else B.table_trunc_linear(L, 8, 32)))
    return out


def run_batch(torch, bt, datas, inits, stream=None):
    dd = [torch.from_numpy(d).cuda() for d in datas]
    ll = [torch.from_numpy(i.copy()).cuda() for i in inits]
    s = torch.cuda.current_stream().cuda_stream if stream is None else stream
    bt.run([t.data_ptr() for t in dd], [t.data_ptr() for t in ll], s)
    torch.cuda.synchronize()
    return dd, ll


@pytest.mark.parametrize("dt,norm,iters", [("float32", 0, 6), ("int32", 0, 5), ("float32", 1, 4), ("float32", 0, 2),
                                           ("float32", 0, 1), ("int32", 0, 0)])
def test_batch_matches_oracle(oracle_mod, dt, norm, iters):
    torch = torch_cuda()
    say(f"batch {dt} norm={norm} iters={iters}")
    insts = small_instances(14, 5 + iters)
    dtype = np.dtype(dt)
    datas = [B.instance_costs(x, dtype) for x in insts]
    rng = np.random.default_rng(1)
    inits = [rng.integers(0, x["L"], size=(x["H"], x["W"])).astype(np.int32) for x in insts]
    bt = B.make_batch(insts, iters=iters, norm=norm, dtype=dtype)
    dd, ll = run_batch(torch, bt, datas, inits)
    for k, x in enumerate(insts):
        ref_lab, ref_msgs, _ = oracle_mod.hbp(datas[k], x["table"].astype(dtype), x["model"], 1, iters, norm=norm,
                                              labels_init=inits[k], return_state=True)
        np.testing.assert_array_equal(ll[k].cpu().numpy(), ref_lab, err_msg=f"instance {k}")
        np.testing.assert_array_equal(bt.dump_msgs(k), ref_msgs, err_msg=f"instance {k}")


def test_batch_tiles_and_replays_identical(oracle_mod):
    torch = torch_cuda()
    say("batch tiles / replays")
    insts = small_instances(12, 77)
    datas = [B.instance_costs(x) for x in insts]
    bt = B.make_batch(insts, iters=7)
    dd = [torch.from_numpy(d).cuda() for d in datas]
    ll = [torch.zeros((x["H"], x["W"]), dtype=torch.int32, device="cuda") for x in insts]
    ptrs = ([t.data_ptr() for t in dd], [t.data_ptr() for t in ll])
    s = torch.cuda.current_stream().cuda_stream
    refs = [oracle_mod.bp(datas[k], x["table"], x["model"], 7) for k, x in enumerate(insts)]
    for tiles, bpc in [((256,) * 4, 4), ((64,) * 4, 2), ((1024, 64, 512, 128), 8), ((128,) * 4, 3)]:
        bt.set_tiles(tiles, bpc)
        for rep in range(3):
            for t in ll:
                t.fill_(-1)
            _ = (dd[0] * 2).sum().item()  # eager work between replays
            bt.run(*ptrs, s)
            torch.cuda.synchronize()
            for k in range(len(insts)):
                lab = ll[k].cpu().numpy()
                inner = lab[1:-1, 1:-1]
                np.testing.assert_array_equal(inner, refs[k][1:-1, 1:-1], err_msg=f"tiles {tiles} rep {rep} inst {k}")


def test_batch_rejects_bad_instances():
    d = [G.make_desc(10, 10, 4, num_levels=2)]
    with pytest.raises(G.GinfError):
        B.Batch(d, [B.table_potts(4, 1)])
    d = [G.make_desc(10, 10, 200)]
    with pytest.raises(G.GinfError):
        B.Batch(d, [B.table_potts(200, 1)])


def test_seeded_batch(oracle_mod):
    """A zero seed reproduces the unseeded batch bit for bit; a random seed
    runs every instance from exactly the injected outboxes."""
    torch = torch_cuda()
    say("seeded batch")
    insts = [x for x in small_instances(10, 31) if x["W"] >= 4 and x["H"] >= 4]
    datas = [B.instance_costs(x) for x in insts]
    iters = 4
    bt = B.make_batch(insts, iters=iters)
    inits = [np.zeros((x["H"], x["W"]), np.int32) for x in insts]
    _, ll_ref = run_batch(torch, bt, datas, inits)
    ref_msgs = [bt.dump_msgs(k) for k in range(len(insts))]
    s = torch.cuda.current_stream().cuda_stream
    preds = []
    for x in insts:
        pw, ph = (x["W"] + 3) // 4, (x["H"] + 3) // 4
        preds.append(torch.zeros((ph, pw, 4 * x["L"]), dtype=torch.float32, device="cuda"))
    bt.set_seeded(True)
    for k, x in enumerate(insts):
        bt.seed(k, preds[k].data_ptr(), preds[k].shape[1], preds[k].shape[0], 0, 1.0, s)
    _, ll = run_batch(torch, bt, datas, inits)
    for k in range(len(insts)):
        np.testing.assert_array_equal(ll[k].cpu().numpy(), ll_ref[k].cpu().numpy(), err_msg=f"zero seed inst {k}")
        np.testing.assert_array_equal(bt.dump_msgs(k), ref_msgs[k], err_msg=f"zero seed inst {k}")
    rng = np.random.default_rng(8)
    for k, x in enumerate(insts):
        preds[k].copy_(torch.from_numpy(rng.standard_normal(tuple(preds[k].shape)).astype(np.float32) * 10))
        bt.seed(k, preds[k].data_ptr(), preds[k].shape[1], preds[k].shape[0], 2, 0.5, s)
    torch.cuda.synchronize()
    seeded = [bt.dump_msgs(k) for k in range(len(insts))]
    _, ll = run_batch(torch, bt, datas, inits)
    for k, x in enumerate(insts):
        m = seeded[k].copy()
        H, W = x["H"], x["W"]
        assert not m[:, :, 0, :].any() and not m[:, :, :, 0].any() and not m[:, :, H - 1, :].any()
        bel = np.zeros_like(datas[k])
        for it in range(iters):
            oracle_mod.half_sweep(datas[k], m, bel, x["table"], x["model"], 0, it)
        want = oracle_mod.labels_from_belief(bel, np.zeros((H, W), np.int32))
        np.testing.assert_array_equal(ll[k].cpu().numpy()[1:-1, 1:-1], want[1:-1, 1:-1], err_msg=f"inst {k}")
        np.testing.assert_array_equal(bt.dump_msgs(k), m, err_msg=f"inst {k}")



@pytest.mark.parametrize("iters,compact", [(2, True), (3, True), (6, True), (4, False)])
def test_batch_predictions_equal_seeding(iters, compact):
    """ginf_batch_set_predictions (the first half-sweep reads each instance's
    prediction) gives the labels and messages of seeding every instance with
    ginf_batch_seed and running seeded, for every label capacity (16-128,
    full and partial), both data-cost forms and repeated replays; switching
    the predictions off returns the unseeded batch."""
    torch = torch_cuda()
    say(f"batch predictions iters={iters} compact={compact}")
    insts = [x for x in small_instances(16, 41) if x["W"] >= 4 and x["H"] >= 4 and x["L"] % 4 == 0]
    assert {x["L"] for x in insts} >= {16, 32, 64, 96, 128}
    datas = [B.instance_costs(x) for x in insts]
    if not compact:
        datas = [d + 0.25 for d in datas]  # fractional costs: 32-bit storage
    inits = [np.zeros((x["H"], x["W"]), np.int32) for x in insts]
    s = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(iters)
    preds = []
    for x in insts:
        pw, ph = (x["W"] + 3) // 4, (x["H"] + 3) // 4
        preds.append(torch.from_numpy(rng.standard_normal((ph, pw, 4 * x["L"])).astype(np.float32) * 8).cuda())
    ref = B.make_batch(insts, iters=iters)
    _, ll_plain = run_batch(torch, ref, datas, inits)
    plain_msgs = [ref.dump_msgs(k) for k in range(len(insts))]
    ref.set_seeded(True)
    for k in range(len(insts)):
        ref.seed(k, preds[k].data_ptr(), preds[k].shape[1], preds[k].shape[0], 2, 0.5, s)
    _, ll_ref = run_batch(torch, ref, datas, inits)
    ref_msgs = [ref.dump_msgs(k) for k in range(len(insts))]
    bt = B.make_batch(insts, iters=iters)
    bt.set_predictions([(preds[k].data_ptr(), preds[k].shape[1], preds[k].shape[0], 0.5) for k in range(len(insts))])
    for rep in range(2):
        _, ll = run_batch(torch, bt, datas, inits)
        for k in range(len(insts)):
            np.testing.assert_array_equal(ll[k].cpu().numpy(), ll_ref[k].cpu().numpy(), err_msg=f"inst {k} rep {rep}")
            np.testing.assert_array_equal(bt.dump_msgs(k), ref_msgs[k], err_msg=f"inst {k} rep {rep}")
    bt.set_predictions(None)
    _, ll = run_batch(torch, bt, datas, inits)
    for k in range(len(insts)):
        np.testing.assert_array_equal(ll[k].cpu().numpy(), ll_plain[k].cpu().numpy(), err_msg=f"off: inst {k}")
        np.testing.assert_array_equal(bt.dump_msgs(k), plain_msgs[k], err_msg=f"off: inst {k}")


def test_batch_predictions_refused():
    """L % 4 != 0, one half-sweep, or int messages: refused with an error."""
    torch = torch_cuda()
    insts = [x for x in small_instances(16, 41) if x["W"] >= 4 and x["H"] >= 4]
    odd = [x for x in insts if x["L"] % 4][:1] + [x for x in insts if x["L"] % 4 == 0][:1]
    assert len(odd) == 2
    p = torch.zeros(1 << 16, dtype=torch.float32, device="cuda")
    for case in ("odd", "iters", "int"):
        xs = odd if case == "odd" else odd[1:]
        bt = B.make_batch(xs, iters=1 if case == "iters" else 3, dtype=np.int32 if case == "int" else np.float32)
        with pytest.raises(G.GinfError):
            bt.set_predictions([(p.data_ptr(), 2, 2, 1.0) for _ in xs])

@pytest.mark.parametrize("compact", [1, 2, 0])
def test_batch_data_storage_forms(oracle_mod, compact):
    """Each instance keeps 8-bit data costs when they fit, else 16-bit ones,
    and falls back on its own when they do not (here: one instance with a
    cost past 255, one fractional and one out-of-range instance among ones
    in [0, 40]); with every setting (narrowest form, at most 16 bits, T)
    every instance equals the oracle."""
    torch = torch_cuda()
    say(f"batch data storage compact={compact}")
    insts = small_instances(12, 123)
    datas = [B.instance_costs(x) for x in insts]
    datas[2] = datas[2] + np.float32(0.25)
    datas[5] = datas[5].copy()
    datas[5][0, 1, 1] = 70000.0
    datas[7] = datas[7].copy()
    datas[7][-1, 2, 1] = 300.0
    bt = B.make_batch(insts, iters=5)
    bt.set_compact_data(compact)
    inits = [np.zeros((x["H"], x["W"]), np.int32) for x in insts]
    dd, ll = run_batch(torch, bt, datas, inits)
    for k, x in enumerate(insts):
        ref_lab, ref_msgs, _ = oracle_mod.hbp(datas[k], x["table"], x["model"], 1, 5, norm=0,
                                              labels_init=inits[k], return_state=True)
        np.testing.assert_array_equal(ll[k].cpu().numpy(), ref_lab, err_msg=f"instance {k}")
        np.testing.assert_array_equal(bt.dump_msgs(k), ref_msgs, err_msg=f"instance {k}")
