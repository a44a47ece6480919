"""The CPU restatement of the learned network's kernels (oracle/learn_oracle.py)
pinned by what the mathematics fixes: a unit-tap kernel shifts the volume,
backward-data is the adjoint of the forward convolution, backward-weights is
the gradient of <conv(x; W), y> in W, and bf16 rounding is round-to-nearest-
even."""
import numpy as np

import learn_oracle as lo

DIMS = (4, 5, 3)   # nx, ny, nz
NV = 60


def _rand(rng, *shape):
    return rng.standard_normal(shape).astype(np.float32)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], np.float32)
    r = lo.to_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_unit_tap_is_a_shift():
    rng = np.random.default_rng(1)
    x = lo.to_bf16(_rand(rng, 1, NV, 16))
    w = np.zeros((16, 27, 16), np.float32)
    t = (1 + 1) + 3 * (0 + 1) + 9 * (-1 + 1)   # dz=+1, dx=0, dy=-1
    w[:, t, :] = np.eye(16)
    _, y = lo.conv_fwd(x, w, np.zeros(16, np.float32), 0, DIMS)
    nx, ny, nz = DIMS
    xg = x.reshape(ny, nx, nz, 16)
    yg = y.reshape(ny, nx, nz, 16)
    for j in range(ny):
        for i in range(nx):
            for k in range(nz):
                src = xg[j - 1, i, k + 1] if (j - 1 >= 0 and k + 1 < nz) else np.zeros(16)
                assert np.array_equal(yg[j, i, k], src)


def test_backward_data_is_adjoint():
    rng = np.random.default_rng(2)
    x = lo.to_bf16(_rand(rng, 2, NV, 16))
    dy = lo.to_bf16(_rand(rng, 2, NV, 16))
    w = lo.to_bf16(_rand(rng, 16, 27, 16))
    _, y = lo.conv_fwd(x, w, np.zeros(16, np.float32), 0, DIMS)
    lhs = float((y.astype(np.float64) * dy).sum())
    # conv_bwd_data rounds to bf16; compare against the unrounded adjoint
    nx, ny, nz = DIMS
    dg = dy.astype(np.float64).reshape(2, ny, nx, nz, 16)
    acc = np.zeros_like(dg)
    for t, dz, dx, dyy in lo.taps():
        acc += lo._shift(dg, -dz, -dx, -dyy) @ w[:, t, :].astype(np.float64)
    rhs = float((acc.reshape(x.shape) * x).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    assert np.abs(lo.conv_bwd_data(dy, w, DIMS) - acc.reshape(x.shape)).max() <= 2 ** -7 * np.abs(acc).max()


def test_backward_weight_is_gradient():
    rng = np.random.default_rng(3)
    x = lo.to_bf16(_rand(rng, 2, NV, 16))
    dy = _rand(rng, 2, NV, 16)
    w = _rand(rng, 16, 27, 16)
    dW, db = lo.conv_bwd_weight(x, dy, DIMS)
    # <conv(x; W) + b, dy> is linear in W and b: its gradient is exact
    h = np.zeros_like(w)
    h[3, 7, 5] = 1.0
    f = lambda W, b: float((lo.conv_fwd(x, W, b, 0, DIMS)[1].astype(np.float64) * dy).sum())
    b0 = np.zeros(16, np.float32)
    assert abs((f(w + h, b0) - f(w, b0)) - dW[3, 7, 5]) <= 1e-3 * max(1.0, abs(dW[3, 7, 5]))
    e = np.zeros(16, np.float32)
    e[4] = 1.0
    assert abs((f(w, e) - f(w, b0)) - db[4]) <= 1e-3 * max(1.0, abs(db[4]))


# ---- the warm start's U-Net (unet_kernels.hip restated in learn_oracle)

def _exact():
    lo.CAST, lo.BF16 = np.float64, False


def _restore():
    lo.CAST, lo.BF16 = np.float32, True


def test_unet_transfers_are_adjoint():
    """interp^T = 8 restrict on node-coincident grids (boundaries included),
    restriction of a constant is the constant in the interior, and
    interpolation reproduces trilinear fields exactly"""
    _exact()
    try:
        dims = (5, 7, 3)
        cd = lo.coarse_dims(dims)
        nf, nc = int(np.prod(dims)), int(np.prod(cd))
        I, R = np.zeros((nf, nc)), np.zeros((nc, nf))
        for j in range(nc):
            e = np.zeros((1, nc, 1)); e[0, j, 0] = 1
            I[:, j] = lo.interp(e, dims)[0, :, 0]
        for j in range(nf):
            e = np.zeros((1, nf, 1)); e[0, j, 0] = 1
            R[:, j] = lo.restrict(e, dims)[0, :, 0]
        assert np.abs(I.T - 8 * R).max() == 0.0
        ones = lo.restrict(np.ones((1, nf, 1)), dims)[0, :, 0].reshape(cd[1], cd[0], cd[2])
        assert np.allclose(ones[1:-1, 1:-1, 1:-1], 1.0)
        # trilinear field a + b x + c y + d z + e x y z on the coarse nodes
        nx, ny, nz = dims
        j, i, k = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
        f = lambda x, y, z: 1.5 + 0.3 * x - 0.7 * y + 0.2 * z + 0.05 * x * y * z  # noqa: E731
        coarse = f(i[::2, ::2, ::2], j[::2, ::2, ::2], k[::2, ::2, ::2]).reshape(1, -1, 1)
        fine = lo.interp(coarse, dims)[0, :, 0]
        assert np.abs(fine - f(i, j, k).reshape(-1)).max() < 1e-12
    finally:
        _restore()


def test_unet_conv_adjoints():
    """with any channel counts: <conv(x), y> = <x, conv^T(y)> and the weight
    gradient is the gradient of the bilinear form <conv(x; W), y> in W"""
    _exact()
    try:
        rng = np.random.default_rng(5)
        dims, nv = (4, 5, 3), 60
        ci, co = 32, 48
        x = rng.standard_normal((2, nv, ci))
        y = rng.standard_normal((2, nv, co))
        w = rng.standard_normal((co, 27, ci))
        _, cx = lo.conv_fwd_c(x, w, None, None, 0, dims)
        assert np.isclose((cx * y).sum(), (x * lo.conv_bwd_data_c(y, w, dims)).sum(), rtol=1e-12)
        dW, db = lo.conv_bwd_weight_c(x, y, dims)
        V = rng.standard_normal(w.shape)
        _, cv = lo.conv_fwd_c(x, V, None, None, 0, dims)
        assert np.isclose((dW * V).sum(), (cv * y).sum(), rtol=1e-12)
        assert np.allclose(db, y.sum(axis=(0, 1)))
    finally:
        _restore()


def test_unet_chain_gradient_by_finite_differences():
    """the whole U-Net's backward composition (the order of the GPU
    training's backward pass) against central differences of <out, g>, in
    float64 without bf16 rounding"""
    _exact()
    try:
        rng = np.random.default_rng(0)
        dims = (17, 17, 17)
        nv = 17 ** 3
        W = [rng.standard_normal((co, 27, ci)) * np.sqrt(2 / (27 * ci)) for ci, co in lo.UNET_LAYERS]
        b = [rng.standard_normal(co) * 0.1 for _, co in lo.UNET_LAYERS]
        x = rng.standard_normal((1, nv, 16))
        a, out, d = lo.unet_forward(x, W, b, dims)
        g = rng.standard_normal(out.shape)
        grads = lo.unet_backward(a, g, W, d)

        def loss(Wk, bk):
            return float((lo.unet_forward(x, Wk, bk, dims)[1] * g).sum())
        eps = 1e-6
        for k, i in ((0, (1, 5, 3)), (6, (40, 20, 7)), (10, (3, 13, 30)), (14, (2, 26, 15))):
            Wp = [w.copy() for w in W]; Wp[k][i] += eps
            Wm = [w.copy() for w in W]; Wm[k][i] -= eps
            fd = (loss(Wp, b) - loss(Wm, b)) / (2 * eps)
            assert abs(fd - grads[k][0][i]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, grads[k][0][i])
        bp = [v.copy() for v in b]; bp[5][2] += eps
        bm = [v.copy() for v in b]; bm[5][2] -= eps
        fd = (loss(W, bp) - loss(W, bm)) / (2 * eps)
        assert abs(fd - grads[5][1][2]) <= 1e-5 * max(1.0, abs(fd))
    finally:
        _restore()


def test_unet_layout_matches_library():
    """the training code's layer table, the restatement's and the library's
    parameter counts agree"""
    import ctypes as C
    from citcom_amd import unet
    assert unet.LAYERS == lo.UNET_LAYERS
    L = unet._lib()
    nw, nb = C.c_long(), C.c_long()
    assert L.cl_unet_params(C.byref(nw), C.byref(nb)) == 0
    assert (nw.value, nb.value) == unet.param_counts() == (594432, 528)
