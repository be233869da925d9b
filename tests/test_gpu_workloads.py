"""The batched benchmark workloads at their real sizes.

Batched plain BP: 64 MRFs (widths log-uniform in [160, 1280], heights in
[120, 720], labels from {16, 32, 64, 128}, Potts w=15 on even indices and
truncated linear lambda=8 tau=4 on odd ones, seeds 3000...), 20 half-sweeps
each, one graph replay of the persistent batched kernels.  Every instance's
labels and level-0 messages equal a single-plan run of that instance (the
reference runs each instance as its own decodeBp, src/bp.cc:182-187), and the
smallest instance of each label class equals the oracle.

Batched learned warm start: one 64-instance chunk of the 256-instance
workload (640x480x64 and 1280x720x96, truncated linear and Potts, the
committed networks, K finest-level half-sweeps), whose labels equal
single-plan seeded runs from the same predictions; every instance that falls
back carries plain HBP's labels and energy exactly, and the flags follow the
0.1 % rule.
"""
import numpy as np
import pytest

from gpu_util import torch_cuda, say
from gridinfcuda_amd import _lib as G
from gridinfcuda_amd import batch as B

pytestmark = pytest.mark.gpu

ITERS = 20
CHUNKS = 4


@pytest.fixture(scope="module")
def batched64():
    torch = torch_cuda()
    say("batched workload: 64 MRFs, 20 half-sweeps (graph replay)")
    from gridinfcuda_amd import autotune
    insts = B.mixed_instances(64, 3000)
    datas = [B.instance_costs(x) for x in insts]
    dd = [torch.from_numpy(d).cuda() for d in datas]
    ll = [torch.zeros((x["H"], x["W"]), dtype=torch.int32, device="cuda") for x in insts]
    bt = B.make_batch(insts, iters=ITERS)
    autotune.apply_cached(bt)
    s = torch.cuda.current_stream().cuda_stream
    bt.run([t.data_ptr() for t in dd], [t.data_ptr() for t in ll], s)
    bt.run([t.data_ptr() for t in dd], [t.data_ptr() for t in ll], s)  # a second replay of the same graph
    torch.cuda.synchronize()
    labels = [t.cpu().numpy() for t in ll]
    yield insts, datas, dd, labels, bt
    bt.close()
    del dd
    torch.cuda.empty_cache()


@pytest.mark.parametrize("chunk", range(CHUNKS))
def test_batched64_equals_single_plans(batched64, chunk):
    torch = torch_cuda()
    insts, datas, dd, labels, bt = batched64
    s = torch.cuda.current_stream().cuda_stream
    per = len(insts) // CHUNKS
    for k in range(chunk * per, (chunk + 1) * per):
        x = insts[k]
        say(f"  instance {k}: {x['W']}x{x['H']}x{x['L']} {'potts' if x['model'] == G.SM_POTTS else 'tl'}")
        desc = G.make_desc(x["W"], x["H"], x["L"], G.F32, x["model"], 1, ITERS, G.NORM_MEAN, 0, 2)
        plan = G.Plan(desc, x["table"])
        lab = torch.zeros((x["H"], x["W"]), dtype=torch.int32, device="cuda")
        plan.run(dd[k].data_ptr(), lab.data_ptr(), s)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(labels[k], lab.cpu().numpy(), err_msg=f"instance {k} labels")
        np.testing.assert_array_equal(bt.dump_msgs(k), plan.dump_msgs(0), err_msg=f"instance {k} messages")
        del plan


def test_batched64_smallest_per_class_equals_oracle(batched64, oracle_mod):
    insts, datas, dd, labels, bt = batched64
    for L in B.LABEL_CLASSES:
        ks = [k for k, x in enumerate(insts) if x["L"] == L]
        assert ks, f"no instance with L={L}"
        k = min(ks, key=lambda i: insts[i]["W"] * insts[i]["H"])
        x = insts[k]
        say(f"  oracle: instance {k} {x['W']}x{x['H']}x{L}")
        ref_lab, ref_msgs, _ = oracle_mod.hbp(datas[k], x["table"], x["model"], 1, ITERS, norm=0, return_state=True,
                                              threads=16)
        np.testing.assert_array_equal(labels[k], ref_lab, err_msg=f"instance {k} labels")
        np.testing.assert_array_equal(bt.dump_msgs(k), ref_msgs, err_msg=f"instance {k} messages")


def test_batched_warm_start_chunk64():
    torch = torch_cuda()
    from gridinfcuda_amd.warmstart_batch import (BatchedWarmStart, class_workloads, fallback_flags, instance_class,
                                                 load_nets)
    K = 14
    say(f"batched warm start: one 64-instance chunk, K={K}")
    classes = class_workloads()
    nets = load_nets(classes)
    ids = list(range(64))
    cc = [instance_class(i) for i in ids]
    assert sorted(set(cc)) == [0, 1, 2, 3]
    bws = BatchedWarmStart(classes, nets, cc, K)
    bws.load([5000 + i for i in ids])
    bws.plain()
    bws.warm(True)
    torch.cuda.synchronize()
    e_plain = bws.energies(bws.plain_labels)
    e_warm = bws.energies(bws.labels)
    warm = [t.cpu().numpy().copy() for t in bws.labels]
    s = torch.cuda.current_stream().cuda_stream
    # every instance's labels equal a single-plan seeded run from the same prediction
    for c, idx in bws.members.items():
        wl = classes[c]
        plan = G.Plan(G.make_desc(wl.W, wl.H, wl.L, G.F32, wl.model, 1, K), wl.table)
        lab = torch.zeros((wl.H, wl.W), dtype=torch.int32, device="cuda")
        for j, k in enumerate(idx):
            lab.zero_()
            plan.seed(0, bws.pred[c][j].data_ptr(), wl.w4, wl.h4, 2, wl.msg_scale, s)
            plan.run_schedule(bws.costs[k].data_ptr(), lab.data_ptr(), 0, [K], True, s)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(warm[k], lab.cpu().numpy(), err_msg=f"instance {k} ({wl.key()})")
        del plan
    flags = bws.finalize(e_warm, e_plain)
    assert flags == fallback_flags(e_warm, e_plain)
    assert all(f == (w > p * 1.001) for f, w, p in zip(flags, e_warm, e_plain))
    say(f"  fallbacks: {sum(flags)} of {len(flags)}; mean excess "
        f"{np.mean([w / p - 1 for w, p in zip(e_warm, e_plain)]):+.4%}")
    e_final = bws.energies(bws.labels)
    for k in ids:
        if flags[k]:
            np.testing.assert_array_equal(bws.labels[k].cpu().numpy(), bws.plain_labels[k].cpu().numpy())
            assert e_final[k] == e_plain[k]
        else:
            np.testing.assert_array_equal(bws.labels[k].cpu().numpy(), warm[k])
            assert e_final[k] <= e_plain[k] * 1.001
    bws.close()
    del bws
    torch.cuda.empty_cache()
