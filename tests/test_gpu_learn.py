"""The learned warm start: its BF16 MFMA kernels against the CPU
restatement (oracle/learn_oracle.py), then training on solutions of the
plain solver and what the trained network gains as the Uzawa initial guess.
"""
import time

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _bf(a):
    import learn_oracle as lo
    return lo.to_bf16(np.asarray(a, np.float32))


def test_kernels_against_oracle():
    import torch
    import learn_oracle as lo
    from citcom_amd import learn
    rng = np.random.default_rng(4)
    dims = (5, 6, 7)   # nx, ny, nz
    nb, nv, CH = 2, 5 * 6 * 7, 16
    x = _bf(rng.standard_normal((nb, nv, CH)))
    w = _bf(rng.standard_normal((CH, 27, CH)) * 0.1)
    bias = rng.standard_normal(CH).astype(np.float32) * 0.1
    net = learn.ConvNet(layers=1)
    net.w[0] = torch.tensor(w).cuda()
    net.wbf[0] = net.w[0].to(torch.bfloat16)
    net.b[0] = torch.tensor(bias).cuda()
    L = learn._cl()
    P = learn._p
    xt = torch.tensor(x).to(torch.bfloat16).cuda()
    for relu in (0, 1):
        y = torch.empty_like(xt)
        yf = torch.empty(xt.shape, dtype=torch.float32, device="cuda")
        assert L.cl_conv3d_fwd(P(xt), nb, *dims, P(net.wbf[0]), P(net.b[0]), relu, P(y), P(yf), None) == 0
        yo_bf, yo = lo.conv_fwd(x, w, bias, relu, dims)
        assert np.abs(yf.cpu().numpy() - yo).max() <= 1e-5 * np.abs(yo).max()
        assert rel_l2(y.float().cpu().numpy(), yo_bf) < 4e-3
    dy = _bf(rng.standard_normal((nb, nv, CH)))
    dyt = torch.tensor(dy).to(torch.bfloat16).cuda()
    dx = torch.empty_like(dyt)
    assert L.cl_conv3d_bwd_data(P(dyt), nb, *dims, P(net.wbf[0]), P(dx), None) == 0
    assert rel_l2(dx.float().cpu().numpy(), lo.conv_bwd_data(dy, w, dims)) < 4e-3
    ws = torch.empty(L.cl_conv3d_bwd_weight_workspace(nb, *dims) // 4, dtype=torch.float32, device="cuda")
    g = torch.empty(CH * 27 * CH + CH, dtype=torch.float32, device="cuda")
    assert L.cl_conv3d_bwd_weight(P(xt), P(dyt), nb, *dims, P(g), P(ws), None) == 0
    dW, db = lo.conv_bwd_weight(x, dy, dims)
    gn = g.cpu().numpy()
    assert rel_l2(gn[:CH * 27 * CH], dW.reshape(-1)) < 1e-5
    assert rel_l2(gn[CH * 27 * CH:], db) < 1e-5
    # ReLU backward and the loss gradient
    yb = torch.tensor(_bf(rng.standard_normal((nb, nv, CH)))).to(torch.bfloat16).cuda()
    d2 = dyt.clone()
    assert L.cl_relu_bwd(P(yb), P(d2), d2.numel(), None) == 0
    assert np.array_equal(d2.float().cpu().numpy(), lo.relu_bwd(yb.float().cpu().numpy(), dy))
    yf = rng.standard_normal((nb, nv, CH)).astype(np.float32)
    tg = rng.standard_normal((nb, nv, CH)).astype(np.float32)
    loss, dyl = net.loss_grad(torch.tensor(yf).cuda(), torch.tensor(tg).cuda())
    lo_loss, lo_dy = lo.mse_grad(yf, tg, learn.NOUT)
    assert loss == pytest.approx(lo_loss / (nb * nv * learn.NOUT), rel=1e-5)
    assert rel_l2(dyl.float().cpu().numpy(), lo_dy) < 4e-3
    say("learned-net kernels match the CPU restatement (fwd, bwd data, bwd weight, relu, mse)")


def test_gradient_through_network():
    """The full forward and backward pass of the 4-layer network on the GPU
    against the same chain composed from the CPU restatement (gradients of
    every layer), and the last layer's gradient against a finite difference
    (the loss is quadratic in the last layer's weights)."""
    import torch
    import learn_oracle as lo
    from citcom_amd import learn
    rng = np.random.default_rng(8)
    dims = (5, 5, 5)
    nb, nv = 2, 125
    net = learn.ConvNet(seed=3)
    x = np.zeros((nb, nv, 16), np.float32)
    x[..., :learn.NIN] = rng.standard_normal((nb, nv, learn.NIN))
    x = _bf(x)
    tg = np.zeros((nb, nv, 16), np.float32)
    tg[..., :learn.NOUT] = rng.standard_normal((nb, nv, learn.NOUT))
    yf = net.forward(torch.tensor(x).to(torch.bfloat16).cuda(), dims)
    _, dy = net.loss_grad(yf, torch.tensor(tg).cuda())
    grads = [g.cpu().numpy() for g in net.backward(dy)]
    W = [_bf(w.cpu().numpy()) for w in net.w]
    B = [b.cpu().numpy() for b in net.b]
    acts = [x]
    for li in range(4):
        a_bf, a_f = lo.conv_fwd(acts[-1], W[li], B[li], li < 3, dims)
        acts.append(a_bf if li < 3 else a_f)
    _, d = lo.mse_grad(acts[4], tg, learn.NOUT)
    for li in reversed(range(4)):
        if li < 3:
            d = lo.relu_bwd(acts[li + 1], d)
        dW, db = lo.conv_bwd_weight(acts[li], d, dims)
        e = rel_l2(grads[li][:16 * 27 * 16], dW.reshape(-1))
        say(f"layer {li}: rel L2 of dW against the CPU chain {e:.2e}")
        assert e < 2e-2
        if li > 0:
            d = lo.conv_bwd_data(d, W[li], dims)

    def loss_last(w3):
        y = lo.conv_fwd(acts[3], w3, B[3], 0, dims)[1].astype(np.float64)
        e = (y - tg)[..., :learn.NOUT]
        return (e * e).sum() / (nb * nv * learn.NOUT)
    g3 = grads[3][:16 * 27 * 16].reshape(16, 27, 16).astype(np.float64)
    h = 1e-3 / max(np.abs(g3).max(), 1e-30)
    fd = (loss_last(W[3] + h * g3) - loss_last(W[3] - h * g3)) / (2 * h)
    an = float((g3 ** 2).sum())
    say(f"last layer: directional derivative analytic {an:.5e}, finite difference {fd:.5e}")
    assert fd == pytest.approx(an, rel=1e-2)


def test_warm_start_gain():
    """Train on the plain solver's solutions (learn16: 16^3, Ra 1e5, random
    temperatures and contrasts 1e2..1e4) and use the network as the Uzawa
    initial guess on held-out problems: the solve reaches the same criterion
    (solution within the solver tolerance of the cold start's) in fewer
    iterations."""
    import torch
    from citcom_amd import learn
    from citcom_amd.hip import Citcom
    t0 = time.time()
    train_set = learn.make_samples("learn16", 96, seed=1)
    test_set = learn.make_samples("learn16", 8, seed=2)
    say(f"samples: {len(train_set)} training + {len(test_set)} test from the plain solver, {time.time() - t0:.1f} s")
    p0 = learn.problem_with_contrast("learn16", 100.0)
    dims = p0.dims(learn.net_level(p0))
    t0 = time.time()
    net, losses = learn.train(train_set, dims, epochs=300, batch=16, lr=2e-3, log=lambda m: say(m))
    say(f"training {time.time() - t0:.1f} s, loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] < 0.1 * losses[0]

    # throughput of the forward convolutions (MFMA)
    xb = torch.zeros((64, dims[0] * dims[1] * dims[2], 16), dtype=torch.bfloat16, device="cuda")
    net.forward(xb, dims)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        net.forward(xb, dims)
    e1.record()
    torch.cuda.synchronize()
    flops = 20 * 4 * 2 * xb.shape[0] * xb.shape[1] * 27 * 16 * 16
    say(f"forward: {flops / (e0.elapsed_time(e1) * 1e-3) / 1e12:.2f} TFLOP/s bf16 "
        f"({flops / (e0.elapsed_time(e1) * 1e-3) / 2.5e15 * 100:.3f} % of the 2.5 PFLOP/s dense MFMA peak; "
        f"{xb.shape[0]} volumes of {xb.shape[1]} nodes)")

    cold_its, warm_its = [], []
    for s in test_set:
        p = learn.problem_with_contrast("learn16", s.contrast)
        c = Citcom(p)
        c.initialize(s.T)
        st = c.general_stokes_solver()
        Uc, Pc = c.download("U"), c.download("P")
        c.close()
        c = Citcom(p)
        c.initialize(s.T)
        learn.warm_start(net, c, p, s.T)
        st_w = c.general_stokes_solver()
        Uw, Pw = c.download("U"), c.download("P")
        c.close()
        cold_its.append(st["iterations"])
        warm_its.append(st_w["iterations"])
        # free-slip walls all round: the pressure is fixed up to a constant,
        # which depends on the initial guess (as in the reference)
        assert rel_l2(Uw, Uc) < 1e-4 and rel_l2(Pw - Pw.mean(), Pc - Pc.mean()) < 1e-4
    say(f"Uzawa iterations on {len(test_set)} held-out problems: cold start {np.mean(cold_its):.1f} "
        f"({cold_its}), learned warm start {np.mean(warm_its):.1f} ({warm_its})")
    assert np.mean(warm_its) < np.mean(cold_its)


def test_coarse_correction_device_matches_host():
    """The coarse correction as the solver runs it (input packing, the four
    MFMA convolutions, scaling back, boundary rows) against the same network
    driven from the host on the same right-hand side."""
    from citcom_amd import deck, learn
    from citcom_amd.hip import Citcom
    from conftest import problem
    p = problem("tiny_tdep")
    c = Citcom(p)
    c.initialize(deck.perturbed_temperature(p))
    c.get_system_viscosity()
    c.construct_stiffness_B_matrix()
    net = learn.ConvNet(seed=3, nin=4, nout=3)
    net.yscale = np.ones(16, np.float32)
    net.yscale[:3] = (0.5, 2.0, 1.5)
    learn.install_coarse_net(net, c, on=False)
    free = learn.coarse_free(c)
    F = np.random.default_rng(5).standard_normal(len(free)) * free
    dev = c.apply("cnet", 0, None, F)
    x, s = learn.coarse_features(F, learn.coarse_logeta(c, p))
    host = learn.coarse_predict(net, x, s, p.dims(0)) * free
    c.close()
    say(f"coarse correction, device vs host network: rel L2 {rel_l2(dev, host):.2e}")
    assert np.all(dev[~free] == 0.0)
    # the log10-viscosity feature is formed in FP32 on the device and FP64 on
    # the host, so its BF16 rounding can differ
    assert rel_l2(dev, host) < 2e-2


def test_coarse_correction_in_multigrid():
    """Train the coarse-level correction on tiny_tdep's family (random
    temperature fields at its viscosity law; target: the coarsest level's
    exact solve), then use it in every V-cycle of the Stokes solve at
    accuracy 1e-9: each inner solve still iterates to its own criterion, so
    the velocity and pressure equal the plain solve's to 1e-6."""
    from citcom_amd import deck, learn
    from citcom_amd.hip import Citcom
    from conftest import problem
    t0 = time.time()
    train_set = learn.make_coarse_samples("tiny_tdep", 3, 24, 16, seed=1)
    test_set = learn.make_coarse_samples("tiny_tdep", 3, 4, 8, seed=2)
    p = problem("tiny_tdep@1e-9")
    dims = p.dims(0)
    net, losses = learn.train_coarse(train_set, dims, epochs=300)
    en, es = learn.coarse_relative_errors(net, test_set, dims)
    say(f"coarse correction: {len(train_set)} samples, loss {losses[0]:.3e} -> {losses[-1]:.3e}, "
        f"{time.time() - t0:.1f} s; held-out error against the exact coarse solve {en:.3f} "
        f"(the {p.v_steps_low}-sweep smoother: {es:.3f})")
    assert losses[-1] < 0.2 * losses[0]
    runs = {}
    for on in (0, 1):
        c = Citcom(p)
        c.initialize(deck.perturbed_temperature(p))
        if on:
            learn.install_coarse_net(net, c)
        st = c.general_stokes_solver()
        runs[on] = (c.download("U"), c.download("P"), st)
        c.close()
    (U0, P0, s0), (U1, P1, s1) = runs[0], runs[1]
    eu, ep = rel_l2(U1, U0), rel_l2(P1 - P1.mean(), P0 - P0.mean())
    say(f"tiny_tdep @1e-9: Uzawa {s0['iterations']} / {s1['iterations']} its, multigrid calls "
        f"{s0['mg_cycles']} (coarse smoother) / {s1['mg_cycles']} (learned correction); U {eu:.2e}, P {ep:.2e}")
    assert eu < 1e-6 and ep < 1e-6


def test_coarse_correction_fallback():
    """The plain multigrid is the learned correction's fallback: an untrained
    network with a huge output scale makes the first multigrid call raise the
    residual; the solve switches the correction off, restarts with the
    coarsest level's smoother and reaches the plain solve's U and P."""
    from citcom_amd import deck, learn
    from citcom_amd.hip import Citcom
    from conftest import problem
    p = problem("tiny_tdep@1e-9")
    net = learn.ConvNet(seed=7, nin=4, nout=3)
    net.yscale = np.full(16, 1e6, np.float32)
    runs = {}
    for on in (0, 1):
        c = Citcom(p)
        c.initialize(deck.perturbed_temperature(p))
        if on:
            learn.install_coarse_net(net, c)
        st = c.general_stokes_solver()
        runs[on] = (c.download("U"), c.download("P"), st, c.coarse_net_state() if on else None)
        c.close()
    (U0, P0, s0, _), (U1, P1, s1, state) = runs[0], runs[1]
    eu, ep = rel_l2(U1, U0), rel_l2(P1 - P1.mean(), P0 - P0.mean())
    say(f"fallback: network state {state}, Uzawa {s0['iterations']} / {s1['iterations']} its; U {eu:.2e}, P {ep:.2e}")
    assert state[0] is False and state[1] >= 1
    # switched off in the first (calibration) solve, the run is the plain
    # run bit for bit
    assert np.array_equal(U1, U0) and np.array_equal(P1, P0)
    assert s1["iterations"] == s0["iterations"] and s1["mg_cycles"] == s0["mg_cycles"]


def test_coarse_correction_layer256():
    """The 256x256x128-element workload (7 levels; two viscosity layers,
    contrast 1e5): the correction trained on the coarse problems of the
    workload family under a 3-level mesh of the same coarse grid.  Its
    first inner solve calibrates it against the coarsest level's smoother:
    either it stays on and the solve takes no more multigrid calls than the
    plain one, or it is switched off and the run is the plain run bit for
    bit.  (piterations 10 bounds the solve; the mixed form.)"""
    from citcom_amd import deck, learn
    from citcom_amd.hip import Citcom
    from conftest import problem
    train_set = learn.make_coarse_samples("layer256", 3, 48, 16, seed=1)
    p = problem("layer256")
    p.p_iterations = 10
    net, losses = learn.train_coarse(train_set, p.dims(0), epochs=400)
    T = deck.seeded_temperature(p, seed=1)
    runs = {}
    for on in (0, 1):
        say(f"layer256 step-0 solve, learned coarse correction {'on' if on else 'off'}")
        c = Citcom(p, mixed=1)
        c.initialize(T)
        if on:
            learn.install_coarse_net(net, c)
        st = c.general_stokes_solver()
        runs[on] = (c.download("U"), c.download("P"), st, c.coarse_net_state() if on else None)
        c.close()
    (U0, P0, s0, _), (U1, P1, s1, state) = runs[0], runs[1]
    say(f"layer256: correction {'kept' if state[0] else 'switched off'} ({state[1]} fallbacks); multigrid calls "
        f"{s0['mg_cycles']} plain / {s1['mg_cycles']} with it; U {rel_l2(U1, U0):.1e}")
    if state[0]:
        assert s1["mg_cycles"] <= s0["mg_cycles"]
        assert rel_l2(U1, U0) < 1e-4
    else:
        assert np.array_equal(U1, U0) and np.array_equal(P1, P0)


def test_coarse_correction_against_oracle():
    """k_cnet (the whole correction in one workgroup) against the numpy
    restatement of its layers (oracle/learn_oracle.py: the same BF16
    operands, products summed in FP64, BF16 activations between layers)."""
    import learn_oracle as lo
    from citcom_amd import deck, learn
    from citcom_amd.hip import Citcom
    from conftest import problem
    p = problem("tiny_tdep")
    c = Citcom(p)
    c.initialize(deck.perturbed_temperature(p))
    c.get_system_viscosity()
    c.construct_stiffness_B_matrix()
    net = learn.ConvNet(seed=11, nin=4, nout=3)
    net.yscale = np.ones(16, np.float32)
    net.yscale[:3] = (1.5, 0.5, 2.0)
    learn.install_coarse_net(net, c, on=False)
    free = learn.coarse_free(c)
    F = np.random.default_rng(9).standard_normal(len(free)) * free
    dev = c.apply("cnet", 0, None, F)
    x, s = learn.coarse_features(F, learn.coarse_logeta(c, p))
    c.close()
    dims = p.dims(0)
    a = lo.to_bf16(x)[None]
    for li in range(net.layers):
        w = net.wbf[li].float().cpu().numpy()
        b = net.b[li].cpu().numpy()
        a_bf, a_f = lo.conv_fwd(a, w, b, int(li < net.layers - 1), dims)
        a = a_bf if li < net.layers - 1 else a_f
    ref = (a[0][:, :3].astype(np.float64) * net.yscale[:3] * s).reshape(-1) * free
    say(f"coarse correction against the oracle: rel L2 {rel_l2(dev, ref):.2e}")
    # FP32 MFMA accumulation against FP64 sums, through BF16 activations
    # (measured 4.4e-8; a BF16 rounding flip of one activation stays far
    # below the bound)
    assert rel_l2(dev, ref) < 1e-4
