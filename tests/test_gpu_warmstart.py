"""GPU tests of the learned warm start: the MFMA convolution kernels against
the float64 oracle (forward, weight and input gradients, a whole network's
gradients), the optimizer, the pooling / target / cost kernels, seeding and
scheduled BP runs, and a short training run that lowers the loss."""
import ctypes

import numpy as np
import pytest

from gpu_util import torch_cuda, say
from gridinfcuda_amd import _lib as G
from oracle import conv_oracle as C

pytestmark = pytest.mark.gpu


def dev_bf16(torch, a):
    return torch.from_numpy(C.to_bf16_bits(a).view(np.int16)).cuda()


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


def rel(a, b):
    return float(np.abs(a - b).max() / max(1e-12, np.abs(b).max()))


@pytest.mark.parametrize("N,H,W,Ci,Co", [(2, 7, 9, 32, 32), (1, 5, 33, 64, 32), (3, 4, 6, 32, 256), (1, 6, 7, 96, 384),
                                         (2, 3, 17, 32, 64), (1, 9, 5, 256, 32), (4, 21, 45, 32, 256),
                                         (2, 37, 70, 64, 32), (1, 5, 11, 160, 32), (1, 4, 9, 192, 64),
                                         (1, 6, 6, 224, 32), (1, 3, 7, 288, 32)])
def test_conv_forward_and_gradients(N, H, W, Ci, Co):
    torch = torch_cuda()
    say(f"conv N={N} H={H} W={W} Ci={Ci} Co={Co}")
    rng = np.random.default_rng(Ci * 7 + Co)
    x = C.bf16_round(rng.standard_normal((N, H, W, Ci)))
    w = C.bf16_round(rng.standard_normal((Co, 3, 3, Ci)) * 0.2)
    b = rng.standard_normal(Co).astype(np.float32)
    xd, wd = dev_bf16(torch, x), dev_bf16(torch, w)
    bd = torch.from_numpy(b).cuda()
    yf = torch.empty((N, H, W, Co), dtype=torch.float32, device="cuda")
    yb = torch.empty((N, H, W, Co), dtype=torch.int16, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    G.check(G.lib().ginf_conv3x3_fwd(vp(xd), vp(wd), vp(bd), N, H, W, Ci, Co, 0, None, vp(yf), s))
    G.check(G.lib().ginf_conv3x3_fwd(vp(xd), vp(wd), vp(bd), N, H, W, Ci, Co, 1, vp(yb), None, s))
    torch.cuda.synchronize()
    ref = C.conv3x3(x, w, b.astype(np.float64))
    assert rel(yf.cpu().numpy(), ref) < 1e-5
    relu_b = C.from_bf16_bits(yb.cpu().numpy().view(np.uint16))
    np.testing.assert_array_equal(relu_b, C.bf16_round(np.maximum(yf.cpu().numpy(), 0)))
    # gradients for dz
    dz = C.bf16_round(rng.standard_normal((N, H, W, Co)))
    dzd = dev_bf16(torch, dz)
    dw = torch.empty((Co, 3, 3, Ci), dtype=torch.float32, device="cuda")
    db = torch.empty(Co, dtype=torch.float32, device="cuda")
    G.check(G.lib().ginf_conv3x3_bwd_weight(vp(xd), vp(dzd), N, H, W, Ci, Co, vp(dw), vp(db), s))
    wf = torch.empty((Ci, 3, 3, Co), dtype=torch.int16, device="cuda")
    G.check(G.lib().ginf_conv3x3_flip_weights(vp(wd), Co, Ci, vp(wf), s))
    dx = torch.empty((N, H, W, Ci), dtype=torch.float32, device="cuda")
    if Co % 32 == 0 and Ci % 16 == 0:
        G.check(G.lib().ginf_conv3x3_fwd(vp(dzd), vp(wf), None, N, H, W, Co, Ci, 0, None, vp(dx), s))
    torch.cuda.synchronize()
    rdx, rdw, rdb = C.conv3x3_bwd(x, w, dz)
    assert rel(dw.cpu().numpy(), rdw) < 1e-5
    assert rel(db.cpu().numpy(), rdb) < 1e-5
    if Co % 32 == 0:
        assert rel(dx.cpu().numpy(), rdx) < 1e-5


@pytest.mark.parametrize("N,H,W,Ci,Co", [(2, 11, 40, 16, 64), (1, 5, 9, 16, 32), (1, 4, 8, 16, 16), (2, 5, 33, 32, 48),
                                         (1, 9, 35, 48, 32), (1, 6, 33, 48, 16), (1, 7, 50, 64, 16),
                                         (1, 6, 37, 128, 16), (1, 5, 20, 96, 48), (1, 3, 5, 80, 32)])
def test_conv_weight_gradient_shapes(N, H, W, Ci, Co):
    """The weight and bias gradients for every (input tiles, co tiles per
    block) instance of the LDS-staged kernel (16..128 input channels, 16..64
    output channels per block) and the per-block fallback (80 inputs)."""
    torch = torch_cuda()
    say(f"conv weight gradient N={N} H={H} W={W} Ci={Ci} Co={Co}")
    rng = np.random.default_rng(Ci * 5 + Co * 3 + W)
    x = C.bf16_round(rng.standard_normal((N, H, W, Ci)))
    w = C.bf16_round(rng.standard_normal((Co, 3, 3, Ci)) * 0.2)
    dz = C.bf16_round(rng.standard_normal((N, H, W, Co)))
    xd, dzd = dev_bf16(torch, x), dev_bf16(torch, dz)
    dw = torch.empty((Co, 3, 3, Ci), dtype=torch.float32, device="cuda")
    db = torch.empty(Co, dtype=torch.float32, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    G.check(G.lib().ginf_conv3x3_bwd_weight(vp(xd), vp(dzd), N, H, W, Ci, Co, vp(dw), vp(db), s))
    torch.cuda.synchronize()
    _, rdw, rdb = C.conv3x3_bwd(x, w, dz)
    assert rel(dw.cpu().numpy(), rdw) < 1e-5
    assert rel(db.cpu().numpy(), rdb) < 1e-5


@pytest.mark.parametrize("L,N,H,W", [(64, 2, 13, 21), (64, 1, 120, 160), (64, 4, 8, 10), (64, 1, 1, 1),
                                     (96, 3, 45, 80), (96, 1, 7, 9), (96, 2, 180, 320), (32, 2, 11, 12)])
def test_infer_matches_forward(L, N, H, W):
    """The one-call network forward (ginf_convnet_fwd: the one-launch fused
    kernel for 64 and 96 labels, layer by layer otherwise) gives the per-layer
    forward's output bit for bit, on image sizes that are and are not
    multiples of its 8 x 10 tile."""
    torch = torch_cuda()
    from gridinfcuda_amd.warmstart import ConvNet
    say(f"one-call network forward L={L} N={N} H={H} W={W}")
    net = ConvNet(L, seed=5)
    rng = np.random.default_rng(6)
    # non-zero biases so that the padding ring of every layer matters
    for b in net.b:
        b.copy_(torch.from_numpy(rng.standard_normal(b.shape).astype(np.float32) * 0.3))
    x = dev_bf16(torch, rng.standard_normal((N, H, W, L)))
    a = net.forward(x, N, H, W).clone()
    b = torch.empty_like(a)
    net.infer(x, N, H, W, b)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_network_gradients_match_oracle():
    """Six-layer net (L=32 -> 32 x5 -> 128) forward + backward on the GPU vs the
    oracle with the same bf16 rounding points."""
    torch = torch_cuda()
    from gridinfcuda_amd.warmstart import ConvNet
    say("network gradient check")
    net = ConvNet(32, seed=3)
    rng = np.random.default_rng(4)
    N, H, W = 2, 6, 8
    x = C.bf16_round(rng.standard_normal((N, H, W, 32)))
    t = rng.standard_normal((N, H, W, 128)).astype(np.float32)
    out = net.forward(dev_bf16(torch, x), N, H, W)
    loss, g = net.mse(torch.from_numpy(t).cuda())
    net.backward(g)
    torch.cuda.synchronize()
    ws = [w.cpu().numpy().astype(np.float64) for w in net.w]
    bs = [b.cpu().numpy().astype(np.float64) for b in net.b]
    rout, ins, outs = C.net_forward(x, ws, bs)
    assert rel(out.cpu().numpy(), rout) < 1e-3
    n = rout.size
    rloss = float(((rout - t) ** 2).sum() / n)
    assert abs(float(loss) - rloss) <= 1e-3 * rloss
    grads, _ = C.net_backward(2 * (rout - t) / n, ws, ins, outs)
    for k in range(6):
        gw = net.gw[k].cpu().numpy()
        gb = net.gb[k].cpu().numpy()
        err_w = np.linalg.norm(gw - grads[k][0]) / max(1e-30, np.linalg.norm(grads[k][0]))
        err_b = np.linalg.norm(gb - grads[k][1]) / max(1e-30, np.linalg.norm(grads[k][1]))
        say(f"  layer {k}: rel grad err w {err_w:.2e} b {err_b:.2e}")
        assert err_w < 2e-2 and err_b < 2e-2


def test_adam_matches_numpy():
    torch = torch_cuda()
    rng = np.random.default_rng(5)
    n = 1000
    w = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    m1 = np.zeros(n, np.float32)
    m2 = np.zeros(n, np.float32)
    wd, gd = torch.from_numpy(w.copy()).cuda(), torch.from_numpy(g).cuda()
    m1d, m2d = torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for step in (1, 2, 3):
        G.check(G.lib().ginf_adam(vp(wd), vp(gd), vp(m1d), vp(m2d), None, n, 1e-2, 0.9, 0.999, 1e-8, step, 0.0, s))
        m1 = 0.9 * m1 + 0.1 * g
        m2 = 0.999 * m2 + 0.001 * g * g
        w = w - 1e-2 * (m1 / (1 - 0.9 ** step)) / (np.sqrt(m2 / (1 - 0.999 ** step)) + 1e-8)
    np.testing.assert_allclose(wd.cpu().numpy(), w, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("W,H,L,dt", [(64, 45, 32, "f32"), (128, 37, 64, "f32"), (64, 44, 20, "f32"),
                                      (70, 45, 32, "f32"), (96, 42, 24, "i32"), (68, 40, 16, "i32")])
def test_minpool_shapes(W, H, L, dt):
    """ginf_minpool_costs against a numpy 4x4 min-pool for every kernel form:
    8 labels per thread (W % 4 == 0, L % 8 == 0), 16-byte rows per label,
    and the per-element fallback; partial blocks at the bottom edge."""
    torch = torch_cuda()
    from gridinfcuda_amd.warmstart import POOL, IN_SCALE
    say(f"minpool W={W} H={H} L={L} {dt}")
    rng = np.random.default_rng(W * H + L)
    c = rng.integers(0, 41, size=(L, H, W))
    costs = torch.from_numpy(c.astype(np.float32 if dt == "f32" else np.int32)).cuda()
    w4, h4 = (W + 3) // 4, (H + 3) // 4
    out = torch.empty((h4, w4, L), dtype=torch.int16, device="cuda")
    G.check(G.lib().ginf_minpool_costs(vp(costs), G.F32 if dt == "f32" else G.I32, W, H, L, POOL,
                                       ctypes.c_float(IN_SCALE), vp(out),
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    pad = np.full((L, h4 * 4, w4 * 4), np.inf)
    pad[:, :H, :W] = c
    ref = pad.reshape(L, h4, 4, w4, 4).min(axis=(2, 4))
    got = C.from_bf16_bits(out.cpu().numpy().view(np.uint16))
    np.testing.assert_array_equal(got, C.bf16_round(np.transpose(ref, (1, 2, 0)).astype(np.float32) * np.float32(IN_SCALE)))



@pytest.mark.parametrize("W,H,L,dt,levels,start", [(64, 45, 32, "f32", 1, 0), (130, 90, 64, "f32", 3, 0),
                                                   (70, 45, 32, "f32", 3, 2), (97, 43, 18, "f32", 2, 1),
                                                   (96, 42, 24, "i32", 2, 0), (67, 41, 96, "i32", 3, 1),
                                                   (50, 30, 300, "f32", 1, 0)])
def test_prepare_minpool_matches(W, H, L, dt, levels, start):
    """ginf_bp_plan_prepare writes the min-pooled network input in the same
    pass as the level data (label-quad layout; the separate kernel for other
    layouts): equal to ginf_minpool_costs, and a run on the prepared data
    (GINF_RUN_PREPARED) equals a run that prepares itself."""
    torch = torch_cuda()
    from gridinfcuda_amd.warmstart import POOL, IN_SCALE
    say(f"prepare + minpool W={W} H={H} L={L} {dt} levels={levels} start={start}")
    rng = np.random.default_rng(W * 7 + H + L)
    c = rng.integers(0, 41, size=(L, H, W))
    if dt == "f32" and L == 64:
        c = c + 0.5  # fractional costs: 32-bit data storage, the relayout's second pass
    np_t = np.float32 if dt == "f32" else np.int32
    costs = torch.from_numpy(c.astype(np_t)).cuda()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    w4, h4 = (W + 3) // 4, (H + 3) // 4
    ref = torch.empty((h4, w4, L), dtype=torch.int16, device="cuda")
    G.check(G.lib().ginf_minpool_costs(vp(costs), G.F32 if dt == "f32" else G.I32, W, H, L, POOL,
                                       ctypes.c_float(IN_SCALE), vp(ref), s))
    a = np.arange(L)
    table = np.minimum(4 * np.abs(a[:, None] - a[None, :]), 12).astype(np_t)
    norm = G.NORM_MIN if L > 256 else G.NORM_MEAN
    plan = G.Plan(G.make_desc(W, H, L, G.F32 if dt == "f32" else G.I32, G.SM_TRUNC_LINEAR, levels, 3, norm), table)
    got = torch.full((h4, w4, L), -1, dtype=torch.int16, device="cuda")
    plan.prepare(costs.data_ptr(), start, got.data_ptr(), IN_SCALE, s.value)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())
    lv = [plan.dump_data(t) for t in range(start + 1)]
    if dt != "f32":
        return  # seeded runs need FP32 messages
    lab = [torch.zeros((H, W), dtype=torch.int32, device="cuda") for _ in range(2)]
    f = 1 << start
    shift = 2 - start
    pw = ((W + f - 1) // f + (1 << shift) - 1) >> shift
    ph = ((H + f - 1) // f + (1 << shift) - 1) >> shift
    pred = torch.from_numpy(rng.standard_normal((ph, pw, 4 * L)).astype(np.float32)).cuda()
    sched = [3] * (start + 1)
    plan.run_predicted(costs.data_ptr(), lab[0].data_ptr(), start, sched, pred.data_ptr(), pw, ph, shift, 0.5,
                       s.value, prepared=True)
    torch.cuda.synchronize()
    m0 = [plan.dump_msgs(t) for t in range(start + 1)]
    for t in range(start + 1):
        np.testing.assert_array_equal(plan.dump_data(t), lv[t])
    plan.run_predicted(costs.data_ptr(), lab[1].data_ptr(), start, sched, pred.data_ptr(), pw, ph, shift, 0.5,
                       s.value)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(lab[0].cpu().numpy(), lab[1].cpu().numpy())
    for t in range(start + 1):
        np.testing.assert_array_equal(plan.dump_msgs(t), m0[t])
    # a level above the prepared one is refused
    if start + 1 < levels:
        with pytest.raises(G.GinfError):
            plan.prepare(costs.data_ptr(), start, 0, 0.0, s.value)
            plan.run_predicted(costs.data_ptr(), lab[0].data_ptr(), start + 1, sched + [3], pred.data_ptr(), pw, ph,
                               shift, 0.5, s.value, prepared=True)

def test_costs_pool_and_targets():
    torch = torch_cuda()
    from gridinfcuda_amd.warmstart import Workload, Pipeline, POOL, IN_SCALE
    say("costs / minpool / block-mean")
    wl = Workload(W=70, H=45, L=32, rects=6, levels=3, iters=4)
    pipe = Pipeline(wl)
    g = G.gen_stereo(11, wl.W, wl.H, wl.L, wl.rects, wl.L - 1, wl.trunc, want=("left", "right", "cost_f32"))
    pipe.load_costs(g["left"], g["right"])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(pipe.costs.cpu().numpy(), g["cost_f32"])
    xin = torch.empty((wl.h4, wl.w4, wl.L), dtype=torch.int16, device="cuda")
    pipe.net_input(xin)
    c = g["cost_f32"]
    ref = np.full((wl.L, wl.h4, wl.w4), np.inf)
    for y in range(wl.H):
        for x in range(wl.W):
            ref[:, y // POOL, x // POOL] = np.minimum(ref[:, y // POOL, x // POOL], c[:, y, x])
    got = C.from_bf16_bits(xin.cpu().numpy().view(np.uint16))
    np.testing.assert_array_equal(got, C.bf16_round(np.transpose(ref, (1, 2, 0)) * IN_SCALE))
    pipe.plain()
    tgt = torch.empty((wl.h4, wl.w4, 4 * wl.L), dtype=torch.float32, device="cuda")
    pipe.target(tgt)
    torch.cuda.synchronize()
    msgs = pipe.plan.dump_msgs(0)  # [l][d][y][x]
    refm = np.zeros((wl.h4, wl.w4, 4 * wl.L))
    for Y in range(wl.h4):
        for X in range(wl.w4):
            blk = msgs[:, :, Y * POOL:(Y + 1) * POOL, X * POOL:(X + 1) * POOL].mean(axis=(2, 3))  # [l][d]
            refm[Y, X] = blk.T.reshape(-1) * wl.msg_scale
    np.testing.assert_allclose(tgt.cpu().numpy(), refm, rtol=1e-5, atol=1e-6)


def test_seed_and_schedule(oracle_mod):
    torch = torch_cuda()
    say("seed / scheduled runs")
    L, W, H = 16, 40, 30
    data = G.gen_stereo(3, W, H, L, 6, 15, 20)["cost_f32"]
    table = oracle_mod.table_trunc_linear(L, 4, 12)
    plan = G.Plan(G.make_desc(W, H, L, G.F32, G.SM_TRUNC_LINEAR, 3, 5), table)
    dd = torch.from_numpy(data).cuda()
    lab = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    # unseeded schedule = plain run
    plan.run(dd.data_ptr(), lab.data_ptr(), s)
    torch.cuda.synchronize()
    ref = lab.cpu().numpy().copy()
    lab.zero_()
    plan.run_schedule(dd.data_ptr(), lab.data_ptr(), 2, [5, 5, 5], False, s)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(lab.cpu().numpy(), ref)
    # a zero seed at the top level is the same as starting from zero messages
    pw, ph = (W + 3) // 4, (H + 3) // 4
    pred = torch.zeros((ph, pw, 4 * L), dtype=torch.float32, device="cuda")
    plan.seed(2, pred.data_ptr(), pw, ph, 0, 1.0, s)
    lab.zero_()
    plan.run_schedule(dd.data_ptr(), lab.data_ptr(), 2, [5, 5, 5], True, s)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(lab.cpu().numpy(), ref)
    # seeded level: interior outboxes = pred / scale, borders 0
    rng = np.random.default_rng(6)
    p = rng.standard_normal((ph, pw, 4 * L)).astype(np.float32)
    pred.copy_(torch.from_numpy(p))
    plan.seed(0, pred.data_ptr(), pw, ph, 2, 0.5, s)
    m = plan.dump_msgs(0)
    for (x, y) in [(1, 1), (W - 2, H - 2), (17, 9), (0, 5), (W - 1, 3)]:
        for d in range(4):
            want = np.zeros(L) if x in (0, W - 1) or y in (0, H - 1) else p[y >> 2, x >> 2, d * L:(d + 1) * L] / 0.5
            np.testing.assert_allclose(m[:, d, y, x], want, rtol=1e-6)
    # seeded level-0 run with k half-sweeps equals the oracle started from those messages
    lab.zero_()
    plan.run_schedule(dd.data_ptr(), lab.data_ptr(), 0, [3], True, s)
    torch.cuda.synchronize()
    msgs0 = m.copy()
    for (x, y) in [(x, y) for y in range(H) for x in range(W) if x in (0, W - 1) or y in (0, H - 1)]:
        msgs0[:, :, y, x] = 0
    bel = np.zeros_like(data)
    for it in range(3):
        oracle_mod.half_sweep(data, msgs0, bel, table, 2, 0, it)
    want = oracle_mod.labels_from_belief(bel, np.zeros((H, W), np.int32))
    np.testing.assert_array_equal(lab.cpu().numpy()[1:-1, 1:-1], want[1:-1, 1:-1])
    np.testing.assert_array_equal(plan.dump_msgs(0), msgs0)



# (W, H, L, levels, start level, iters from the start level down to 0, integer costs)
# 640x480 at 16 and 32 labels has enough tiles for the fused pairs of
# half-sweeps (on by default there), so every arrangement of unfused first
# half-sweeps and pairs at the predicted level is covered (2-9 half-sweeps).
PRED_CASES = [
    (640, 480, 16, 1, 0, (2,), True), (640, 480, 16, 1, 0, (3,), True), (640, 480, 16, 1, 0, (4,), True),
    (640, 480, 16, 1, 0, (5,), True), (640, 480, 16, 1, 0, (6,), True), (640, 480, 16, 1, 0, (7,), True),
    (640, 480, 16, 1, 0, (8,), True), (640, 480, 16, 1, 0, (9,), False), (640, 480, 32, 2, 1, (6, 9), True),
    (640, 480, 32, 2, 0, (7,), False), (300, 200, 16, 1, 0, (5,), True), (97, 61, 12, 3, 2, (4, 3, 5), True),
    (130, 90, 64, 3, 0, (5,), True), (130, 90, 64, 3, 1, (4, 6), False), (130, 90, 64, 3, 2, (2, 2, 3), True),
    (83, 77, 96, 2, 0, (4,), False), (70, 50, 128, 2, 1, (3, 2), True),
    # fallbacks to seeding first: L % 4 != 0, one half-sweep at the seeded level, chunked layout (min norm, int-valued)
    (90, 70, 18, 1, 0, (4,), True), (90, 70, 16, 2, 1, (1, 4), True), (90, 70, 300, 1, 0, (3,), True),
]


@pytest.mark.parametrize("W,H,L,levels,start,iters,integer", PRED_CASES)
def test_run_predicted_equals_seed_then_run(W, H, L, levels, start, iters, integer):
    """ginf_bp_plan_run_predicted (the first half-sweep reads the prediction)
    gives the labels and messages of ginf_bp_plan_seed + a seeded schedule."""
    torch = torch_cuda()
    say(f"run_predicted {W}x{H}x{L} levels={levels} start={start} iters={iters} integer={integer}")
    rng = np.random.default_rng(W * 31 + L)
    data = G.gen_stereo(11, W, H, L, 6, L - 1, 20)["cost_f32"]
    if not integer:
        data = data + rng.random(data.shape, dtype=np.float32) * 0.25
    other = G.gen_stereo(12, W, H, L, 6, L - 1, 20)["cost_f32"]
    a = np.arange(L)
    table = np.minimum(5 * np.abs(a[:, None] - a[None, :]), 15).astype(np.float32)
    norm = G.NORM_MIN if L > 256 else G.NORM_MEAN
    desc = G.make_desc(W, H, L, G.F32, G.SM_TRUNC_LINEAR, levels, 5, norm)
    shift = 2 - start
    f = 1 << start
    lw, lh = (W + f - 1) // f, (H + f - 1) // f
    pw, ph = (lw + (1 << shift) - 1) >> shift, (lh + (1 << shift) - 1) >> shift
    pred = torch.from_numpy(rng.standard_normal((ph, pw, 4 * L)).astype(np.float32)).cuda()
    dd = torch.from_numpy(data).cuda()
    do = torch.from_numpy(other).cuda()
    s = torch.cuda.current_stream().cuda_stream
    sched = list(reversed(iters))
    outs = []
    for direct in (False, True):
        plan = G.Plan(desc, table)
        if W == 640:
            assert plan.pairs()  # the fused pairs run at these sizes
        lab = torch.zeros((H, W), dtype=torch.int32, device="cuda")
        plan.run(do.data_ptr(), lab.data_ptr(), s)  # leave another problem's messages in every array
        lab.fill_(-7)
        if direct:
            plan.run_predicted(dd.data_ptr(), lab.data_ptr(), start, sched, pred.data_ptr(), pw, ph, shift, 0.5, s)
        else:
            plan.seed(start, pred.data_ptr(), pw, ph, shift, 0.5, s)
            plan.run_schedule(dd.data_ptr(), lab.data_ptr(), start, sched, True, s)
        torch.cuda.synchronize()
        outs.append((lab.cpu().numpy(), [plan.dump_msgs(t) for t in range(start + 1)]))
        plan.close()
    np.testing.assert_array_equal(outs[1][0], outs[0][0])
    for t in range(start + 1):
        np.testing.assert_array_equal(outs[1][1][t], outs[0][1][t], err_msg=f"level {t}")

def test_training_lowers_loss():
    torch = torch_cuda()
    from gridinfcuda_amd.warmstart import Workload, train
    say("short training run")
    wl = Workload(W=96, H=64, L=32, rects=8, levels=3, iters=6)
    net, hist = train(wl, n_instances=96, batch=16, lr=3e-3, reuse=4, seed_base=50, log=say)
    first, last = hist[0][1], hist[-1][1]
    say(f"  loss {first:.4f} -> {last:.4f}")
    assert last < 0.7 * first


def _halfsweeps_to_plain_energy(G, wl, pipe, ws, lab, e_plain, ks):
    """Fewest level-0 half-sweeps k in ks after seeding level 0 (from the
    network, or from zeros) whose energy is within 0.1 % of plain HBP's;
    None when no k in ks gets there."""
    s = pipe.stream()
    for k in ks:
        lab.zero_()
        ws.run(pipe.costs, lab, [k], seed_level=0)
        e = G.energy_f64(wl.W, wl.H, wl.L, G.F32, pipe.costs.data_ptr(), pipe.table_dev.data_ptr(), lab.data_ptr(), s)
        if e <= e_plain * 1.001:
            return k, e
    return None, None


def test_trained_net_needs_fewer_halfsweeps_than_zero_seed():
    """Ablation on the warm-start claim itself: with the committed 640x480x64
    truncated-linear network, seeding the finest level from its prediction
    reaches plain HBP's energy (within 0.1 %) in fewer level-0 half-sweeps than
    the same seeding from zero messages, on each of three instances."""
    torch = torch_cuda()
    from gridinfcuda_amd.warmstart import ConvNet, Pipeline, WarmStart, Workload, default_model_path
    say("trained network vs zero seed: level-0 half-sweeps to plain energy")
    wl = Workload()
    net = ConvNet.load(default_model_path(wl))
    pipe = Pipeline(wl)
    ws_net, ws_zero = WarmStart(wl, net), WarmStart(wl, None)
    lab = torch.zeros((wl.H, wl.W), dtype=torch.int32, device="cuda")
    ks = list(range(2, 41, 2))
    for seed in (2_000_000, 2_000_001, 2_000_002):
        g = G.gen_stereo(seed, wl.W, wl.H, wl.L, wl.rects, wl.L - 1, wl.trunc, want=("left", "right"))
        pipe.load_costs(g["left"], g["right"])
        pipe.plain()
        e_plain = pipe.energy()
        k_net, e_net = _halfsweeps_to_plain_energy(G, wl, pipe, ws_net, lab, e_plain, ks)
        k_zero, e_zero = _halfsweeps_to_plain_energy(G, wl, pipe, ws_zero, lab, e_plain, ks)
        say(f"  seed {seed}: plain E={e_plain:.0f}; net {k_net} half-sweeps (E={e_net}), "
            f"zero seed {k_zero} half-sweeps (E={e_zero})")
        assert k_net is not None
        assert k_zero is None or k_net < k_zero


def test_batched_warm_start_pipeline():
    """Batched inference + predicted batched BP (the first half-sweep reads
    the predictions): every instance's labels equal a single-plan seeded run
    from the same prediction, and the fallback takes the plain HBP labels
    exactly when the energy rule says so."""
    torch = torch_cuda()
    from gridinfcuda_amd.warmstart import ConvNet, Workload
    from gridinfcuda_amd.warmstart_batch import BatchedWarmStart, fallback_flags
    say("batched warm start pipeline")
    classes = [Workload(W=96, H=64, L=32, model=G.SM_TRUNC_LINEAR, lam=4, tau=3, rects=6, levels=3, iters=5),
               Workload(W=70, H=50, L=64, model=G.SM_POTTS, lam=15, tau=1, rects=4, levels=3, iters=5)]
    nets = [ConvNet(32, seed=1), ConvNet(64, seed=2)]
    cc = [0, 1, 1, 0, 1]
    K = 6
    bws = BatchedWarmStart(classes, nets, cc, K)
    bws.load([70 + k for k in range(len(cc))])
    bws.plain()
    bws.warm(True)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream().cuda_stream
    for c, idx in bws.members.items():
        wl = classes[c]
        plan = G.Plan(G.make_desc(wl.W, wl.H, wl.L, G.F32, wl.model, 1, K), wl.table)
        for j, k in enumerate(idx):
            lab = torch.zeros((wl.H, wl.W), dtype=torch.int32, device="cuda")
            plan.seed(0, bws.pred[c][j].data_ptr(), wl.w4, wl.h4, 2, wl.msg_scale, s)
            plan.run_schedule(bws.costs[k].data_ptr(), lab.data_ptr(), 0, [K], True, s)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(bws.labels[k].cpu().numpy(), lab.cpu().numpy(), err_msg=f"instance {k}")
    e_plain = bws.energies(bws.plain_labels)
    e_warm = bws.energies(bws.labels)
    warm = [t.cpu().numpy().copy() for t in bws.labels]
    flags = bws.finalize(e_warm, e_plain)
    assert flags == fallback_flags(e_warm, e_plain)
    for k in range(len(cc)):
        want = bws.plain_labels[k].cpu().numpy() if flags[k] else warm[k]
        np.testing.assert_array_equal(bws.labels[k].cpu().numpy(), want)
    assert fallback_flags([1.0011, 1.0009], [1.0, 1.0]) == [True, False]
    # A deliberately bad prediction (every seed message scaled far out of
    # range) must fall back: each flagged instance then carries plain HBP's
    # labels and energy exactly.
    for c, idx in bws.members.items():  # the batch's first half-sweep reads these predictions
        wl = classes[c]
        bws.nets[c].forward(bws.xin[c], len(idx), wl.h4, wl.w4, out=bws.pred[c])
        bws.pred[c].mul_(-1e4)
    bws.batch.run([t.data_ptr() for t in bws.costs], [t.data_ptr() for t in bws.labels], s)
    torch.cuda.synchronize()
    e_bad = bws.energies(bws.labels)
    flags = bws.finalize(e_bad, e_plain)
    say(f"  bad prediction: fallback flags {flags}")
    assert any(flags)
    e_final = bws.energies(bws.labels)
    for k in range(len(cc)):
        if flags[k]:
            np.testing.assert_array_equal(bws.labels[k].cpu().numpy(), bws.plain_labels[k].cpu().numpy())
            assert e_final[k] == e_plain[k]
    bws.close()
