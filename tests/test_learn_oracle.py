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
