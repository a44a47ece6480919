"""The learned warm start's U-Net on the GPU: every kernel against the CPU
restatement (oracle/learn_oracle.py), the library's forward pass against
the training code's composition of the same kernels, the whole backward
pass against the restatement's chain (itself pinned by finite differences
in tests/test_learn_oracle.py), and the warm start inside the solver: an
untrained network still ends each solve on the plain solve's solution.
"""
import numpy as np
import pytest

from conftest import problem, rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _bf(a):
    import learn_oracle as lo
    return lo.to_bf16(np.asarray(a, np.float32))


def _t(a, dtype=None):
    import torch
    t = torch.tensor(np.asarray(a, np.float32))
    return (t.to(dtype) if dtype is not None else t).cuda()


def test_unet_kernels_against_oracle():
    import torch
    import learn_oracle as lo
    from citcom_amd import unet
    from citcom_amd.learn import _p
    L = unet._lib()
    rng = np.random.default_rng(3)
    dims = (5, 7, 9)
    nv, nb, ci, co = 5 * 7 * 9, 2, 32, 48
    say("unet kernels: conv fwd/bwd, transfers, loss at", dims)
    x = _bf(rng.standard_normal((nb, nv, ci)))
    w = _bf(rng.standard_normal((co, 27, ci)) * 0.05)
    bias = (rng.standard_normal(co) * 0.1).astype(np.float32)
    res = _bf(rng.standard_normal((nb, nv, co)))
    xt, wt_, bt, rt = _t(x, torch.bfloat16), _t(w, torch.bfloat16), _t(bias), _t(res, torch.bfloat16)
    for relu, with_res in ((1, True), (0, False)):
        y = torch.empty((nb, nv, co), dtype=torch.bfloat16, device="cuda")
        yf = torch.empty((nb, nv, co), dtype=torch.float32, device="cuda")
        assert L.cl_conv3d_fwd_c(_p(xt), nb, *dims, ci, co, _p(wt_), _p(bt), _p(rt) if with_res else None, relu, 0,
                                 _p(y), _p(yf), None) == 0
        yo_bf, yo = lo.conv_fwd_c(x, w, bias, res if with_res else None, relu, dims)
        assert np.abs(yf.cpu().numpy() - yo).max() <= 1e-5 * np.abs(yo).max()
        assert rel_l2(y.float().cpu().numpy(), yo_bf) < 4e-3
    dy = _bf(rng.standard_normal((nb, nv, co)))
    dyt = _t(dy, torch.bfloat16)
    wT = torch.empty((ci, 27, co), dtype=torch.bfloat16, device="cuda")
    assert L.cl_conv3d_transpose_w(_p(wt_), ci, co, _p(wT), None) == 0
    assert np.array_equal(wT.float().cpu().numpy(), w.transpose(2, 1, 0))
    dx = torch.empty((nb, nv, ci), dtype=torch.bfloat16, device="cuda")
    assert L.cl_conv3d_bwd_data_c(_p(dyt), nb, *dims, ci, co, _p(wT), _p(dx), None) == 0
    assert rel_l2(dx.float().cpu().numpy(), lo.conv_bwd_data_c(dy, w, dims)) < 4e-3
    ws = torch.empty(L.cl_conv3d_bwd_weight_c_workspace(nb, *dims, ci, co) // 4, dtype=torch.float32, device="cuda")
    g = torch.empty(co * 27 * ci + co, dtype=torch.float32, device="cuda")
    assert L.cl_conv3d_bwd_weight_c(_p(xt), _p(dyt), nb, *dims, ci, co, _p(g), _p(ws), None) == 0
    dW, db = lo.conv_bwd_weight_c(x, dy, dims)
    gn = g.cpu().numpy()
    assert rel_l2(gn[:co * 27 * ci], dW.reshape(-1)) < 1e-5
    assert rel_l2(gn[co * 27 * ci:], db) < 1e-5
    # transfers with a scale, both directions
    cd = lo.coarse_dims(dims)
    ncv = int(np.prod(cd))
    rc = torch.empty((nb, ncv, ci), dtype=torch.bfloat16, device="cuda")
    assert L.cl_restrict(_p(xt), nb, *dims, ci, 8.0, _p(rc), None) == 0
    assert rel_l2(rc.float().cpu().numpy(), lo.restrict(x, dims, 8.0)) < 4e-3
    xc = _bf(rng.standard_normal((nb, ncv, ci)))
    fi = torch.empty((nb, nv, ci), dtype=torch.bfloat16, device="cuda")
    assert L.cl_interp(_p(_t(xc, torch.bfloat16)), nb, *dims, ci, 0.125, _p(fi), None) == 0
    assert rel_l2(fi.float().cpu().numpy(), lo.interp(xc, dims, 0.125)) < 4e-3
    # loss over 4 of 16 channels with channel weights
    yf = rng.standard_normal((nb, nv, 16)).astype(np.float32)
    tg = rng.standard_normal((nb, nv, 4)).astype(np.float32)
    wch = np.array([1.0, 1.0, 1.0, 2.0], np.float32)
    net_loss, dyl = unet.UNet.loss_grad(None, _t(yf), _t(tg), _t(wch))
    lo_loss, lo_dy = lo.wmse_grad(yf, 16, tg, 4, wch)
    assert net_loss == pytest.approx(lo_loss / (nb * nv * 4), rel=1e-5)
    assert rel_l2(dyl.float().cpu().numpy(), lo_dy) < 4e-3


def test_unet_forward_backward_against_oracle():
    """the training composition (forward, backward) at 17^3 nodes against the
    restatement's chain with the same bf16 rounding points; the library's
    own forward pass (cl_unet_forward) gives the training forward's output
    bit for bit"""
    import torch
    import learn_oracle as lo
    from citcom_amd import unet
    from citcom_amd.learn import _p
    say("unet forward/backward at 17^3")
    rng = np.random.default_rng(7)
    dims = (17, 17, 17)
    nv = 17 ** 3
    net = unet.UNet(seed=2)
    for k in range(len(unet.LAYERS)):   # biases nonzero, output layer O(1)
        net.b[k] = torch.tensor(rng.standard_normal(unet.LAYERS[k][1]) * 0.1, dtype=torch.float32).cuda()
    net.w[-1] = net.w[-1] * 10
    net.wbf = [w.to(torch.bfloat16) for w in net.w]
    x = _bf(rng.standard_normal((1, nv, 16)))
    xt = _t(x, torch.bfloat16)
    out = net.forward(xt, dims)
    W = [w.float().cpu().numpy() for w in net.wbf]
    B = [b.cpu().numpy() for b in net.b]
    a, out_o, d = lo.unet_forward(x, W, B, dims)
    assert rel_l2(out.cpu().numpy(), out_o) < 2e-2   # bf16 activations, FP32 sums in another order
    # the library's forward pass (no activations kept) against the composition
    L = unet._lib()
    wblob, bblob = net.blob()
    wd = torch.from_numpy(wblob.view(np.int16)).cuda()
    bd = torch.from_numpy(bblob).cuda()
    ws = torch.empty(L.cl_unet_workspace(1, *dims), dtype=torch.uint8, device="cuda")
    out2 = torch.empty_like(out)
    assert L.cl_unet_forward(_p(xt), 1, *dims, _p(wd), _p(bd), _p(out2), _p(ws), None) == 0
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.int32), out2.view(torch.int32))
    # backward from the same output gradient, against the restatement fed
    # the GPU's own activations (so each layer's gradient is compared on the
    # same inputs)
    g = _bf(rng.standard_normal((1, nv, 16)) * 1e-3)
    grads = net.backward(_t(g, torch.bfloat16))
    acts = {k: v.float().cpu().numpy() for k, v in net.acts.items()}
    go = lo.unet_backward(acts, g, W, d)
    for k, (ci, co) in enumerate(unet.LAYERS):
        gk = grads[k].cpu().numpy()
        e_w = rel_l2(gk[:co * 27 * ci], go[k][0].reshape(-1))
        e_b = rel_l2(gk[co * 27 * ci:], go[k][1])
        assert e_w < 2e-2 and e_b < 2e-2, (k, e_w, e_b)
    say("unet gradients of all 15 layers agree with the restatement")


def test_untrained_warm_start_keeps_the_solution():
    """the learned warm start with a random network: each solve still ends on
    the plain solve's solution (the guess only moves where the iteration
    starts); the features and target records are finite"""
    import torch
    from citcom_amd import deck, unet
    from citcom_amd.hip import Citcom
    name = "learn16@1e-8"
    p = problem(name)
    say("untrained warm start on", name)
    T = deck.seeded_temperature(p, seed=3)
    runs = {}
    nno = int(np.prod(p.dims()))
    feats = torch.zeros((1, nno, 16), dtype=torch.bfloat16, device="cuda")
    targs = torch.zeros((1, nno, 4), dtype=torch.float32, device="cuda")
    for mode in (0, 2):
        c = Citcom(p, mixed=1)
        if mode:
            w, b = unet.UNet(seed=5).blob()
            c.set_warm_net(w, b)
            c.set_option("warm_start", 2)
        c.initialize(T)
        c.general_stokes_solver()
        its = []
        for s in range(4):
            if mode and s == 3:
                c.warm_record(feats.data_ptr(), targs.data_ptr())
            its.append(c.timestep()["stokes"]["iterations"])
        if mode:
            sc = c.warm_scales()
            assert sc[7] == 1.0 and np.all(np.isfinite(sc)) and sc[3] > 0 and sc[4] > 0
            c.warm_record(None, None)
        runs[mode] = (c.download("U"), c.download("P"), c.download("T"), its)
        c.close()
    say("iterations plain start", runs[0][3], "learned start", runs[2][3])
    assert rel_l2(runs[2][0], runs[0][0]) < 1e-6
    P0, P2 = runs[0][1], runs[2][1]
    assert rel_l2(P2 - P2.mean(), P0 - P0.mean()) < 1e-5
    assert rel_l2(runs[2][2], runs[0][2]) < 1e-6
    assert torch.isfinite(feats.float()).all() and torch.isfinite(targs).all()
    assert feats.float().abs().sum() > 0 and targs.abs().sum() > 0
